"""fdvgpu_motion_search_noalloc (the allocation-free variant) against
fdvgpu_motion_search: identical tables on two calls in a row, the second on a
smaller frame (the shared scratch keeps the first call's contents), with and
without the proposer; sizes beyond its scratch fail with an error."""
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


def run(cur, r0, r1, w, h, navail, qp, weights, noalloc):
    cy = dev(cur[:w * h])
    b0, off, rs = L.padded(np.ascontiguousarray(r0[:w * h]), w, h)
    b1, _, _ = L.padded(np.ascontiguousarray((r1 if r1 is not None else r0)[:w * h]), w, h)
    p0, p1 = dev(b0), dev(b1)
    table = torch.full((((w // 16) * (h // 16)) * G.MOTION_DTYPE.itemsize,), 0xAB, dtype=torch.uint8, device="cuda:0")
    wd = dev(weights) if weights is not None else None
    stream = torch.cuda.current_stream().cuda_stream
    G.motion_search(cy.data_ptr(), p0.data_ptr() + off, p1.data_ptr() + off, rs, w, h, navail, qp, table.data_ptr(),
                    wd.data_ptr() if wd is not None else 0, 0.7, stream, noalloc=noalloc)
    torch.cuda.synchronize()
    return table.cpu().numpy().tobytes()


@pytest.mark.parametrize("learned", [False, True])
def test_noalloc_matches_the_allocating_entry_point(learned):
    W = G.load_proposer() if learned else None
    big = content.scene_pan(640, 368, 91, nobj=6, smin=16, smax=96)
    small = content.scene_objects(160, 96, 92, nobj=3, smin=8, smax=40, vmax=10, grain=2.0)
    for k, (sc, navail) in enumerate(((big, 2), (small, 1))):
        w, h = sc.w, sc.h
        f0, f1, f2 = sc.frame(0), sc.frame(1), sc.frame(2)
        say(f"  no-alloc motion search, call {k + 1}: {w}x{h}, navail {navail}, proposer {learned}")
        a = run(f2, f1, f0 if navail == 2 else None, w, h, navail, 26, W, noalloc=False)
        b = run(f2, f1, f0 if navail == 2 else None, w, h, navail, 26, W, noalloc=True)
        assert a == b, f"call {k + 1} ({w}x{h}): tables differ"


def test_noalloc_rejects_frames_beyond_its_scratch():
    t = torch.zeros(64, dtype=torch.uint8, device="cuda:0")
    p = t.data_ptr()
    with pytest.raises(G.GpuError, match="scratch holds"):
        G.motion_search(p, p, p, 8192 + 64, 8192, 4096, 1, 24, p, noalloc=True)
