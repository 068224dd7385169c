"""A caller of the reference's API moved to the device path by one include:
oracle/dropin_demo.c includes the reference's fdv.h, then
include/fdv_amd_dropin.h, and calls fdv_image_encode / fdv_image_decode,
fdv_video_encode / fdv_video_decode, fdv_vtile_encode / fdv_vtile_decode and
the streaming fdv_enc_* / fdv_dec_* unchanged (fdv.h:871-876, 903-908,
973-981, 1126-1198), key frames through the library's own coders.  The
demonstration checks, in the same process, that the output is byte-identical to
the reference's own functions' and that both decoders return the same frames.
It is built by oracle/Makefile into oracle/_ref/ where the reference tree is
present.
"""
import os
import subprocess

import pytest

from gpucases import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "oracle", "_ref", "dropin_demo")


@pytest.mark.parametrize("w,h,n,qp,keyint,bmr", [(320, 192, 12, 28, 5, 4), (640, 368, 8, 24, 0, 8)])
def test_unchanged_call_sites_give_the_references_streams(w, h, n, qp, keyint, bmr):
    assert os.path.exists(DEMO), "oracle/_ref/dropin_demo is built by oracle/Makefile"
    say(f"  drop-in demonstration {w}x{h}, {n} frames, QP {qp}, keyint {keyint}, band {bmr} MB rows")
    r = subprocess.run([DEMO, str(w), str(h), str(n), str(qp), str(keyint), str(bmr)],
                       capture_output=True, text=True, timeout=120)
    say(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dropin ok" in r.stdout
