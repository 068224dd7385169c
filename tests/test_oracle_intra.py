"""Pins the oracle's key-frame codec (orc_image_encode, orc_iframe_encode)
before the device path is compared with it.

1. Golden digests written by the reference build (tests/golden/intra_vectors.json,
   by tests/golden/make_intra_golden.py): bytes and reconstruction of 16 planes
   whose containers cover all three entropy modes.  Run everywhere.
2. Where the reference build is present (oracle/_ref): seeded planes of random
   size, content and QP, and whole three-plane key frames, byte for byte.
3. Closed forms: the nine 4x4 predictors on constant edges, the vertical /
   horizontal / DC definitions, and H.264's transposition symmetries
   (diagonal-down-right is its own transpose with top and left exchanged,
   vertical-right is horizontal-down's); a flat plane at the predictor's own
   DC value codes as one 64x64 leaf per unit with no levels; units that cross
   the plane's edge split without a flag.
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import intraplanes as IP  # noqa: E402
import orclib as L  # noqa: E402

with open(os.path.join(HERE, "golden", "intra_vectors.json")) as f:
    FIX = json.load(f)["cases"]


@pytest.mark.parametrize("case", FIX, ids=lambda c: "%s-%dx%d-qp%d" % tuple(c["case"][:4]))
def test_image_matches_reference_digests(case):
    kind, w, h, qp, seed = case["case"]
    out, rec = L.image_encode(L.oracle(), "orc_image_encode", IP.intra_plane(kind, w, h, seed), qp)
    assert len(out) == case["len"]
    assert int(out[5]) == case["container"]
    assert hashlib.sha256(out.tobytes()).hexdigest() == case["bytes_sha256"]
    assert hashlib.sha256(rec.tobytes()).hexdigest() == case["rec_sha256"]


def test_fixture_covers_every_container():
    assert sorted({c["container"] for c in FIX}) == [0, 1, 2]


needs_ref = pytest.mark.skipif(L.ref() is None, reason="reference build (oracle/_ref) not present")


@needs_ref
def test_image_matches_reference_build_random():
    r = np.random.default_rng(2024)
    for i in range(24):
        w, h = int(r.integers(1, 48)) * 4, int(r.integers(1, 36)) * 4
        qp = int(r.integers(0, 52))
        kind = IP.KINDS[i % len(IP.KINDS)]
        src = IP.intra_plane(kind, w, h, 100 + i)
        a, ra = L.image_encode(L.ref(), "fdv_image_encode", src, qp)
        b, rb = L.image_encode(L.oracle(), "orc_image_encode", src, qp)
        assert a.tobytes() == b.tobytes(), (kind, w, h, qp)
        assert (ra == rb).all(), (kind, w, h, qp)


@needs_ref
@pytest.mark.parametrize("w,h,qp", [(96, 64, 22), (200, 120, 35), (64, 48, 8)])
def test_iframe_matches_reference_build(w, h, qp):
    y = IP.intra_plane("mixed", w, h, w + qp)
    u = IP.intra_plane("ramp", w // 2, h // 2, 3)
    v = IP.intra_plane("sine", w // 2, h // 2, 4)
    outs, recs = [], []
    for lib, fn in ((L.ref(), "fdvref_iframe_encode"), (L.oracle(), "orc_iframe_encode")):
        cap = w * h * 8 + 65536
        out = np.zeros(cap, np.uint8)
        ry = np.zeros((h, w), np.uint8)
        ru = np.zeros((h // 2, w // 2), np.uint8)
        rv = np.zeros_like(ru)
        n = getattr(lib, fn)(y.ctypes.data, u.ctypes.data, v.ctypes.data, w, h, qp, out.ctypes.data, cap,
                             ry.ctypes.data, ru.ctypes.data, rv.ctypes.data)
        assert n > 0
        outs.append(out[:n].tobytes())
        recs.append((ry, ru, rv))
    assert outs[0] == outs[1]
    for a, b in zip(*recs):
        assert (a == b).all()


def _pred4(mode, top, left, tl, ht=1, hl=1):
    p = np.zeros(16, np.uint8)
    top = np.ascontiguousarray(top, np.uint8)
    left = np.ascontiguousarray(left, np.uint8)
    L.oracle().orc_intra4(mode, top.ctypes.data, left.ctypes.data, tl, ht, hl, p.ctypes.data)
    return p.reshape(4, 4).astype(int)


def test_intra4_closed_forms():
    for c in (0, 77, 255):
        for m in range(9):
            assert (_pred4(m, [c] * 8, [c] * 4, c) == c).all()
    r = np.random.default_rng(7)
    for _ in range(200):
        top, left, tl = r.integers(0, 256, 8), r.integers(0, 256, 4), int(r.integers(0, 256))
        assert (_pred4(1, top, left, tl) == np.tile(top[:4], (4, 1))).all()
        assert (_pred4(2, top, left, tl) == np.tile(left[:, None], (1, 4))).all()
        assert (_pred4(0, top, left, tl) == (top[:4].sum() + left.sum() + 4) >> 3).all()
        assert (_pred4(0, top, left, tl, 1, 0) == (top[:4].sum() + 2) >> 2).all()
        assert (_pred4(0, top, left, tl, 0, 1) == (left.sum() + 2) >> 2).all()
        assert (_pred4(0, top, left, tl, 0, 0) == 128).all()
        # exchanging top and left transposes diagonal-down-right, and turns
        # vertical-right into horizontal-down (and back)
        swap_t = np.concatenate([left, left[3:].repeat(4)])
        assert (_pred4(4, swap_t, top[:4], tl) == _pred4(4, top, left, tl).T).all()
        assert (_pred4(6, swap_t, top[:4], tl) == _pred4(5, top, left, tl).T).all()
        assert (_pred4(5, swap_t, top[:4], tl) == _pred4(6, top, left, tl).T).all()


def test_flat_plane_is_one_leaf_per_unit():
    src = np.full((128, 192), 128, np.uint8)
    t = L.image_tree(src, 24)
    assert (t["split0"] == 0).all()
    assert (t["nmode0"] == 0).all()
    assert (t["lev"] == 0).all()
    assert (t["rec"] == 128).all()
    assert (t["rflag"] == 0).all()


def test_units_crossing_the_edge_split_without_a_flag():
    t = L.image_tree(IP.intra_plane("mixed", 72, 40, 2), 30)
    assert (t["split0"] == 2).all()                        # every 64x64 unit crosses the edge
    assert (t["split1"][:, 2] == 2).all()                  # 32x32 nodes at x = 64 cross it too
    out, _ = L.image_encode(L.oracle(), "orc_image_encode", IP.intra_plane("mixed", 72, 40, 2), 30)
    assert out[6] == 0                                     # no depth-0 split flag was sent
