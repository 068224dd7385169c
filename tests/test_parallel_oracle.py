"""Pinning of the oracle's motion-table and block-parallel functions
(oracle/fdv_oracle.c: orc_metab_exhaustive, orc_metab_candidates,
orc_pframe_analyse_table) against functions the reference's vectors and brute
force already pin:

- the motion table is orc_me_search's exhaustive search of every block;
- the candidate search equals it wherever the proposals cover the whole-pel
  minimum, falls back to it below the confidence threshold, and flags SKIP by
  its own threshold;
- the MB loop run on the exhaustive table is the exhaustive mode's MB loop
  (orc_pframe_analyse with the exhaustive search): the same decisions, streams
  and reconstruction, so the block-parallel modes, which reproduce the table
  loop, reproduce the exhaustive mode;
- MBs flagged SKIP-likely are decided between SKIP and INTER16 only;
- the payloads of every table decode with the reference decoder to the
  encoder's reconstruction.
"""
import ctypes as C

import numpy as np
import pytest

import orclib as L
from fdv_amd import content

PROPOSAL = np.dtype([("cx", "i1", 8), ("cy", "i1", 8), ("kx", "i1", 4), ("ky", "i1", 4),
                     ("qx", "i1", (4, 2)), ("qy", "i1", (4, 2)), ("conf", "<f4"), ("pskip", "<f4")])


def _luma(f, w, h):
    return np.ascontiguousarray(f[:w * h]).reshape(h, w)


def _search(cur, ref, w, h, bx, by, bs):
    lib = L.oracle()
    buf, off, st = L.padded(_luma(ref, w, h), w, h)
    mx, my = C.c_int(), C.c_int()
    s = lib.orc_me_search(C.c_void_p(np.ascontiguousarray(cur[:w * h]).ctypes.data), w,
                          C.c_void_p(buf.ctypes.data + off), st, bx, by, bs, bs, 16, 0, 0, 1,
                          C.byref(mx), C.byref(my))
    return mx.value, my.value, s


def _wholepel_min(cur, ref, w, h, bx, by, bs=16):
    """Brute-force whole-pel minimum over +-16 (ties: |dx|+|dy|, then raster)."""
    P = np.pad(_luma(ref, w, h), 32, mode="edge").astype(int)
    blk = _luma(cur, w, h)[by:by + bs, bx:bx + bs].astype(int)
    best = None
    for dy in range(-16, 17):
        for dx in range(-16, 17):
            s = int(np.abs(blk - P[32 + by + dy:32 + bs + by + dy, 32 + bx + dx:32 + bs + bx + dx]).sum())
            key = (s, abs(dx) + abs(dy), (dy + 16) * 33 + dx + 16)
            if best is None or key < best[0]:
                best = (key, dx, dy)
    return best[1], best[2]


def _frames():
    sc = content.scene_objects(96, 64, 71, nobj=3, smin=8, smax=32, vmax=6, grain=1.0)
    return sc.w, sc.h, [sc.frame(i) for i in range(3)]


def test_table_is_the_exhaustive_search_of_every_block():
    w, h, (f0, f1, f2) = _frames()
    for navail, r0, r1 in ((2, f1, f0), (1, f1, None)):
        mt = L.metab_to_np(L.oracle_metab_exhaustive(f2, r0, r1, w, h, navail))
        for i in range(len(mt)):
            bx, by = (i % (w // 16)) * 16, (i // (w // 16)) * 16
            for r, ref in enumerate((r0, r1 if navail == 2 else r0)):
                mx, my, s = _search(f2, ref, w, h, bx, by, 16)
                assert (int(mt[i]["mv16"][r][0]), int(mt[i]["mv16"][r][1]), int(mt[i]["sad16"][r])) == (mx, my, s)
            for q in range(4):
                mx, my, s = _search(f2, r0, w, h, bx + (q & 1) * 8, by + (q >> 1) * 8, 8)
                assert (int(mt[i]["mv8"][q][0]), int(mt[i]["mv8"][q][1]), int(mt[i]["sad8"][q])) == (mx, my, s)
            assert mt[i]["flags"] == 0


def test_candidate_search_covers_the_exhaustive_minimum_and_thresholds():
    w, h, (f0, f1, f2) = _frames()
    nmb = (w // 16) * (h // 16)
    full = L.metab_to_np(L.oracle_metab_exhaustive(f2, f1, None, w, h, 1))
    # below the confidence threshold: the exhaustive table, flags by the SKIP threshold
    low = np.zeros(nmb, PROPOSAL)
    low["conf"] = 0.5
    low["pskip"] = np.where(np.arange(nmb) % 2 == 0, 0.9, 0.1)
    mt = L.metab_to_np(L.oracle_metab_candidates(f2, f1, None, w, h, 1, low, thr=0.7, sthr=0.7))
    assert np.array_equal(mt["mv16"], full["mv16"]) and np.array_equal(mt["sad16"], full["sad16"])
    assert np.array_equal(mt["mv8"], full["mv8"])
    assert list(mt["flags"]) == [1 if i % 2 == 0 else 0 for i in range(nmb)]
    # confident proposals whose first vector's 3x3 neighbourhood holds the whole-pel
    # minimum: the same 16x16 vector and SAD as the full search
    hi = np.zeros(nmb, PROPOSAL)
    hi["conf"] = 0.9
    for i in range(nmb):
        bx, by = (i % (w // 16)) * 16, (i // (w // 16)) * 16
        dx, dy = _wholepel_min(f2, f1, w, h, bx, by)
        hi["cx"][i] = dx - (dx & 1)              # even, as the proposer's grid; +-1 covers dx
        hi["cy"][i] = dy - (dy & 1)
        hi["cx"][i][1:] = 8                      # far-away decoys
        hi["cy"][i][1:] = -8
    mt = L.metab_to_np(L.oracle_metab_candidates(f2, f1, None, w, h, 1, hi, thr=0.7, sthr=0.7))
    assert np.array_equal(mt["mv16"][:, 0], full["mv16"][:, 0])
    assert np.array_equal(mt["sad16"][:, 0], full["sad16"][:, 0])
    assert not mt["flags"].any()
    # the same with the minimum held only by the third correlation-map position
    co = np.zeros(nmb, PROPOSAL)
    co["conf"] = 0.9
    co["cx"], co["cy"] = 8, -8
    co["kx"], co["ky"] = -10, 10
    co["kx"][:, 2], co["ky"][:, 2] = hi["cx"][:, 0], hi["cy"][:, 0]
    mt = L.metab_to_np(L.oracle_metab_candidates(f2, f1, None, w, h, 1, co, thr=0.7, sthr=0.7))
    assert np.array_equal(mt["mv16"][:, 0], full["mv16"][:, 0])
    assert np.array_equal(mt["sad16"][:, 0], full["sad16"][:, 0])
    # the quadrants' minima held only by the quadrant candidates: the same 8x8
    # vectors and SADs as the full search
    qc = np.zeros(nmb, PROPOSAL)
    qc["conf"] = 0.9
    qc["cx"], qc["cy"], qc["kx"], qc["ky"] = 8, -8, -10, 10
    qc["qx"], qc["qy"] = 10, 10
    for i in range(nmb):
        bx, by = (i % (w // 16)) * 16, (i // (w // 16)) * 16
        for q in range(4):
            dx, dy = _wholepel_min(f2, f1, w, h, bx + (q & 1) * 8, by + (q >> 1) * 8, 8)
            qc["qx"][i][q][1], qc["qy"][i][q][1] = dx - (dx & 1), dy - (dy & 1)
    mt = L.metab_to_np(L.oracle_metab_candidates(f2, f1, None, w, h, 1, qc, thr=0.7, sthr=0.7))
    assert np.array_equal(mt["mv8"], full["mv8"]) and np.array_equal(mt["sad8"], full["sad8"])
    # and with every candidate a decoy, the table is a different one
    co["kx"][:, 2], co["ky"][:, 2] = -10, 10
    mt = L.metab_to_np(L.oracle_metab_candidates(f2, f1, None, w, h, 1, co, thr=0.7, sthr=0.7))
    assert not np.array_equal(mt["mv16"][:, 0], full["mv16"][:, 0])


@pytest.mark.parametrize("qp", [20, 28, 36])
def test_table_loop_on_the_exhaustive_table_is_the_exhaustive_mode(qp):
    """orc_pframe_analyse_table takes its vectors from the table; the exhaustive
    mode runs orc_me_search inside the loop.  The exhaustive search ignores the
    predictor, so the two must agree on every decision, stream byte and sample,
    over a chain whose later frames use both references."""
    sc = content.scene_objects(128, 64, 73, nobj=4, smin=8, smax=40, vmax=8, grain=2.0)
    w, h = sc.w, sc.h
    frames = [sc.frame(i) for i in range(4)]
    near, far = frames[0], None
    seen = set()
    for f in range(1, 4):
        navail = 1 if f == 1 else 2
        st, rec, info = L.oracle_analyse(frames[f], near, far, w, h, navail, qp, me_mode=1)
        pst, prec, pinfo = L.oracle_parallel(frames[f], near, far, w, h, navail, qp)
        for i in range(len(info)):
            a, g = info[i], pinfo[i]
            assert (a.mode, a.ref, a.merge_idx, a.sub) == (g.mode, g.ref, g.merge_idx, g.sub), f"frame {f} MB {i}"
            assert [tuple(a.mv[k]) for k in range(4)] == [tuple(g.mv[k]) for k in range(4)], f"frame {f} MB {i}"
            seen.add(a.mode)
        assert pst == st and np.array_equal(prec, rec), f"frame {f}"
        far, near = near, rec
    assert {0, 1, 2}.issubset(seen)


def test_skip_flagged_mbs_weigh_skip_against_inter16_only():
    """A table MB with flags bit 0 set is decided between SKIP and INTER16: with
    every MB flagged no MB is INTRA or INTER8, and the first MB, whose predictor
    and merge list depend on no other decision, keeps an unflagged SKIP or
    INTER16 decision (the INTRA and INTER8 costs it drops did not win)."""
    sc = content.scene_blobs(128, 64, 94, nblob=40, vmax=14, grain=3.0, cut=9)
    w, h = sc.w, sc.h
    f0, f1, f2 = sc.frame(0), sc.frame(1), sc.frame(2)
    qp = 30
    mt = L.oracle_metab_exhaustive(f2, f1, f0, w, h, 2)
    _, _, info = L.oracle_parallel(f2, f1, f0, w, h, 2, qp, metab=mt)
    assert {m.mode for m in info} & {2, 3}, "content gives no INTRA or INTER8 decision to prune"
    for i in range(len(info)):
        mt[i].flags = 1
    _, _, pinfo = L.oracle_parallel(f2, f1, f0, w, h, 2, qp, metab=mt)
    assert all(m.mode in (0, 1) for m in pinfo)
    # the first MB's predictor and merge list do not depend on other decisions
    if info[0].mode in (0, 1):
        assert (pinfo[0].mode, pinfo[0].ref) == (info[0].mode, info[0].ref)


need_ref = pytest.mark.skipif(L.ref() is None, reason="reference build not present (oracle/_ref)")


def _proposals_from_table(mt, seed):
    """Confident proposals around the table's vectors with random decoys: the
    candidate search then mixes hits and misses, as the learned mode does."""
    rng = np.random.default_rng(seed)
    m = L.metab_to_np(mt)
    p = np.zeros(len(m), PROPOSAL)
    p["conf"] = np.where(rng.random(len(m)) < 0.8, 0.9, 0.3)
    p["pskip"] = rng.random(len(m))
    for i in range(len(m)):
        dx, dy = int(m["mv16"][i][0][0]) >> 2, int(m["mv16"][i][0][1]) >> 2
        p["cx"][i] = rng.integers(-8, 9, 8) * 2
        p["cy"][i] = rng.integers(-8, 9, 8) * 2
        if rng.random() < 0.7:
            p["cx"][i][0], p["cy"][i][0] = dx - (dx & 1), dy - (dy & 1)
        p["kx"][i], p["ky"][i] = rng.integers(-8, 9, 4) * 2, rng.integers(-8, 9, 4) * 2
        p["qx"][i], p["qy"][i] = rng.integers(-8, 9, (4, 2)) * 2, rng.integers(-8, 9, (4, 2)) * 2
    return p


@need_ref
@pytest.mark.parametrize("table", ["exhaustive", "candidates", "candidates+skip"])
def test_block_parallel_payloads_decode_with_the_reference_decoder(table):
    """The parallel and learned modes only change the encoder's choices: their
    payloads, packed with the carried entropy state, decode with the reference
    decoder (fdv.h:6492) to exactly the encoder's reconstruction, over a chain
    of P-frames whose later frames reference both earlier ones."""
    scenes = [content.scene_objects(96, 64, 91, nobj=4, smin=8, smax=32, vmax=10, grain=2.0),
              content.scene_blobs(128, 64, 92, nblob=40, vmax=14, grain=3.0, cut=3)]
    sthr = 0.6 if table == "candidates+skip" else 2.0
    for sc in scenes:
        w, h = sc.w, sc.h
        frames = [sc.frame(i) for i in range(5)]
        tce, tcd = L.TabCache(), L.TabCache()
        near, far = frames[0], None
        for f in range(1, 5):
            navail = 1 if f == 1 else 2
            qp = (22, 30, 38, 26)[f - 1]
            mt = L.oracle_metab_exhaustive(frames[f], near, far, w, h, navail)
            if table != "exhaustive":
                mt = L.oracle_metab_candidates(frames[f], near, far, w, h, navail,
                                               _proposals_from_table(mt, f), 0.7, sthr)
            st, rec, info = L.oracle_parallel(frames[f], near, far, w, h, navail, qp, metab=mt)
            pay = L.oracle_pack(qp, st, tce)
            dec = L.ref_pframe_decode(pay, near, far, w, h, navail, tcd)
            assert dec is not None, f"{w}x{h} frame {f}: the reference decoder rejected the payload"
            assert np.array_equal(dec, rec), f"{w}x{h} frame {f}: decoded frame differs from the encoder's recon"
            assert np.array_equal(tce.buf, tcd.buf), f"{w}x{h} frame {f}: entropy state"
            far, near = near, rec
