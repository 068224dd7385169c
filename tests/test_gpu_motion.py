"""fdvgpu_motion_search: the motion table of a whole frame on caller-owned
device buffers equals the oracle's exhaustive table (orc_metab_exhaustive,
itself pinned to the per-block search of fdv_me_search's exhaustive variant),
and with the proposer's weights equals the learned mode's table of a context
coding the same frame."""
import numpy as np
import pytest
import torch

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G
from gpucases import say

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def padded_luma(frame, w, h):
    buf, off, stride = L.padded(np.ascontiguousarray(frame[:w * h]), w, h)
    return dev(buf), off, stride


def search(cur, r0, r1, w, h, navail, qp, weights=None, conf=0.7):
    cy = dev(cur[:w * h])
    p0, off, rs = padded_luma(r0, w, h)
    p1, _, _ = padded_luma(r1 if r1 is not None else r0, w, h)
    nmb = (w // 16) * (h // 16)
    table = torch.zeros(nmb * G.MOTION_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    wd = dev(weights) if weights is not None else None
    G.motion_search(cy.data_ptr(), p0.data_ptr() + off, p1.data_ptr() + off, rs, w, h, navail, qp, table.data_ptr(),
                    wd.data_ptr() if wd is not None else 0, conf)
    torch.cuda.synchronize()
    return np.frombuffer(table.cpu().numpy().tobytes(), dtype=G.MOTION_DTYPE)


@pytest.mark.parametrize("size", [(96, 64), (320, 192)])
def test_exhaustive_table_matches_oracle(size):
    w, h = size
    sc = content.scene_objects(w, h, 81, nobj=4, smin=8, smax=48, vmax=10, grain=2.0)
    f0, f1, f2 = sc.frame(0), sc.frame(1), sc.frame(2)
    for navail, r0, r1 in ((1, f1, None), (2, f1, f0)):
        say(f"  motion search {w}x{h} navail {navail}")
        got = search(f2, r0, r1, w, h, navail, 24)
        want = L.metab_to_np(L.oracle_metab_exhaustive(f2, r0, r1, w, h, navail))
        assert got.tobytes() == want.tobytes()


def test_learned_table_matches_the_context_path():
    sc = content.scene_pan(320, 192, 82, nobj=4, smin=16, smax=64)
    w, h = sc.w, sc.h
    f0, f1 = sc.frame(0), sc.frame(1)
    W = G.load_proposer()
    an = G.Analyser([(w, h)], nsrc=1)
    an.set_proposer(W)
    an.put_ref(0, 0, f0)
    an.put_source(0, 0, f1)
    an.pframe([0], [28], navail=[1], me_mode=G.ME_LEARNED)
    an.fetch()
    want = an.motion(0)
    an.close()
    say("  motion search with the proposer, 320x192")
    got = search(f1, f0, None, w, h, 1, 28, weights=W)
    assert got.tobytes() == want.tobytes()


def test_bad_arguments_fail():
    with pytest.raises(G.GpuError):
        G.motion_search(0, 0, 0, 128, 64, 48, 1, 24, 0)
