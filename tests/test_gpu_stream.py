"""The streaming surface on the device path against the reference's own
(oracle/_ref), key frames from the reference's I-frame coder in both:

- fdvamd_enc_* writes the bytes fdv_enc_* writes: at a fixed QP and under
  average-bitrate control (whose requantising retry reruns decisions and
  coding on the device's kept motion tables), single-stream and tiled, with
  held (all-SKIP) frames, one of them on a key-frame slot;
- fdvamd_dec_* returns the frames fdv_dec_* returns, frame by frame and after
  seeks backwards, forwards within a key-frame interval and across one;
- fdvamd_read / fdvamd_decode agree with fdv_read / fdv_decode.
"""
import ctypes as C

import numpy as np
import pytest

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G
from gpucases import say
from streamlib import Streaming

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def libs():
    assert L.ref() is not None, "the reference build (oracle/_ref) is needed"
    enc, dec = L.ref_iframe_coder()
    G.set_iframe_coder(enc, dec)
    G.set_sequence_mode(G.ME_DIAMOND)
    yield Streaming(G.lib(), "fdvamd_"), Streaming(L.ref(), "fdv_")
    G.set_iframe_coder(0, 0)


def frames_of(w, h, n, seed):
    sc = content.scene_pan(w, h, seed, nobj=3, smin=8, smax=32, pan=(2, -1), grain=2.0)
    return [sc.frame(i) for i in range(n)]


CASES = [
    ("single, fixed QP", None, 0),
    ("single, 120 kbps", None, 120000),
    ("tiled 4 rows, fixed QP", (4, 4), 0),
    ("tiled 4 rows, 120 kbps", (4, 4), 120000),
]


@pytest.mark.parametrize("name,tiled,bps", CASES, ids=[c[0] for c in CASES])
def test_streaming_encoder_writes_the_references_bytes(libs, name, tiled, bps):
    ours, ref = libs
    w, h = 320, 192
    fr = frames_of(w, h, 14, 71)
    say(f"  {name}: {w}x{h}, 14 frames + 2 held, keyint 5")
    a = ours.encode(fr, w, h, 28, 5, tiled=tiled, bitrate=bps, hold=(3, 5))
    b = ref.encode(fr, w, h, 28, 5, tiled=tiled, bitrate=bps, hold=(3, 5))
    say(f"  {len(a)} B vs the reference's {len(b)} B")
    assert a == b


def test_streaming_720p_eight_bands_under_bitrate_control(libs):
    """The 1280x720 eight-band layout at 1 Mbps: one QP for the whole picture
    from the rate control, every band's P-frame in one device pass."""
    ours, ref = libs
    fr = frames_of(1280, 720, 10, 72)
    say("  1280x720, 8 bands, 10 frames, 1 Mbps")
    a = ours.encode(fr, 1280, 720, 24, 60, tiled=(6, 8), bitrate=1000000)
    b = ref.encode(fr, 1280, 720, 24, 60, tiled=(6, 8), bitrate=1000000)
    assert a == b


def test_bitrate_control_changes_the_stream(libs):
    """Under a budget this tight the rate control moves the quantizer: the
    file differs from the fixed-QP one and is less than half its size."""
    ours, _ = libs
    fr = frames_of(320, 192, 8, 73)
    fixed = ours.encode(fr, 320, 192, 20, 0)
    tight = ours.encode(fr, 320, 192, 20, 0, bitrate=60000)
    assert tight != fixed and len(tight) < len(fixed) / 2


@pytest.mark.parametrize("tiled", [None, (4, 4)], ids=["single", "tiled"])
def test_streaming_decoder_returns_the_references_frames(libs, tiled):
    ours, ref = libs
    w, h = 320, 192
    fr = frames_of(w, h, 16, 74)
    data = ref.encode(fr, w, h, 30, 6, tiled=tiled, hold=(4,))
    da, db = ours.decoder(data), ref.decoder(data)
    assert da.fs == db.fs == w * h * 3 // 2
    n = da.info.nframes
    assert n == 17
    for k in range(n):
        ra, xa = da.next()
        rb, xb = db.next()
        assert ra == rb == 1 and np.array_equal(xa, xb), f"frame {k}"
    assert da.next()[0] == db.next()[0] == 0
    for t in (15, 2, 7, 8, 12, 0, 16):
        say(f"  seek {t}")
        assert da.seek(t) == 0 and db.seek(t) == 0
        assert da.pos() == db.pos() == t
        ra, xa = da.next()
        rb, xb = db.next()
        assert ra == rb == 1 and np.array_equal(xa, xb), f"after seek {t}"
    assert da.seek(n) == db.seek(n) == -1
    da.close()
    db.close()


def test_container_read_and_decode(libs):
    ours, ref = libs
    w, h = 320, 192
    fr = frames_of(w, h, 7, 75)
    for tiled in (None, (4, 2)):
        data = ref.encode(fr, w, h, 26, 4, tiled=tiled)
        ra, ia, arr_a = ours.read(data)
        rb, ib, arr_b = ref.read(data)
        assert ra == rb == 0
        for f in ("kind", "version", "fps", "w", "h", "nframes", "qp", "keyint", "nbands", "payload_len"):
            assert getattr(ia, f) == getattr(ib, f), f
        oa = np.zeros(ia.nframes * w * h * 3 // 2, np.uint8)
        ob = np.zeros_like(oa)
        assert ours.f("decode")(C.byref(ia), oa.ctypes.data, 4) == 0
        assert ref.f("decode")(C.byref(ib), ob.ctypes.data, 4) == 0
        assert np.array_equal(oa, ob)
    # malformed heads are refused alike
    bad = bytearray(data)
    bad[4] = 9
    assert ours.read(bytes(bad))[0] == ref.read(bytes(bad))[0] == -1
