"""The host half of the P-frame decoder on the CPU (no device needed):

- fdvamd_unpack_pframe inverts the entropy stage: the oracle's payloads, packed
  with the entropy state carried over a chain of P-frames, unpack to exactly
  the streams that were packed, the unpacker's entropy state follows the
  packer's, and all four entropy codings (one table, a table per stream, the
  cached table, adaptive streams) occur;
- fdvamd_parse_pframe reads the decisions back: every MB's mode, reference,
  vectors, merge index and intra sub-mode equal the oracle encoder's;
- malformed payloads (truncated, corrupted) are rejected or parsed without
  fault, never read out of bounds.
"""
import numpy as np
import pytest

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G


def _chain(sc, nfr, qps):
    """Oracle streams, payloads (entropy state carried) and decisions."""
    w, h = sc.w, sc.h
    frames = [sc.frame(i) for i in range(nfr)]
    tce = L.TabCache()
    near, far = frames[0], None
    out = []
    for f in range(1, nfr):
        navail = 1 if f == 1 else 2
        qp = qps[(f - 1) % len(qps)]
        st, rec, info = L.oracle_analyse(frames[f], near, far, w, h, navail, qp)
        pay = L.oracle_pack(qp, st, tce)
        out.append((qp, st, pay, info, tce.buf.copy()))
        far, near = near, rec
    return out


def test_unpack_inverts_the_entropy_stage_in_every_mode():
    modes = set()
    for sc, qps in ((content.scene_pan(160, 96, 41, nobj=4, smin=16, smax=48), [20, 28, 36]),
                    (content.scene_objects(128, 64, 42, nobj=3, smin=8, smax=32, vmax=6), [30]),
                    (content.scene_blobs(96, 64, 43, nblob=20, vmax=10, grain=2.0, cut=5), [24, 40])):
        tcd = G.TabCache()
        for qp, st, pay, info, tce_after in _chain(sc, 9, qps):
            modes.add(pay[1])
            q, got = G.unpack_pframe(pay, (sc.w // 16) * (sc.h // 16), tcd)
            assert q == qp
            assert got == st
            assert np.array_equal(tcd.buf, tce_after)
    assert modes == {0, 1, 2, 3}, f"entropy codings seen: {sorted(modes)}"


def test_parse_reads_back_every_decision():
    for sc in (content.scene_pan(160, 96, 44, nobj=4, smin=16, smax=48),
               content.scene_blobs(128, 64, 45, nblob=30, vmax=14, grain=3.0, cut=4)):
        seen = set()
        for qp, st, pay, info, _ in _chain(sc, 6, [22, 34]):
            codes = G.parse_pframe(st, sc.w, sc.h)
            for i, m in enumerate(info):
                c = codes[i]
                assert c["mode"] == m.mode, f"MB {i}"
                seen.add(m.mode)
                if m.mode in (0, 1):
                    assert (c["ref"], c["mv"][0][0], c["mv"][0][1]) == (m.ref, m.mv[0][0], m.mv[0][1]), f"MB {i}"
                if m.mode == 0:
                    assert c["merge_idx"] == m.merge_idx and c["ckind"] == 0
                if m.mode == 2:
                    assert c["sub"] == m.sub and c["ckind"] == 2
                if m.mode == 3:
                    assert [tuple(c["mv"][k]) for k in range(4)] == [tuple(m.mv[k]) for k in range(4)]
        assert {0, 1, 2} <= seen


def test_malformed_payloads_are_rejected_without_fault():
    sc = content.scene_pan(96, 64, 46, nobj=3, smin=8, smax=32)
    nmb = (sc.w // 16) * (sc.h // 16)
    chain = _chain(sc, 4, [26])
    rng = np.random.default_rng(7)
    rejected = 0
    for qp, st, pay, info, _ in chain:
        for cut in (0, 1, 2, len(pay) // 2, len(pay) - 1):
            with pytest.raises(G.GpuError):
                G.unpack_pframe(pay[:cut], nmb)
        for trial in range(200):
            b = bytearray(pay)
            for _ in range(1 + trial % 4):
                b[int(rng.integers(0, len(b)))] ^= int(rng.integers(1, 256))
            try:
                _, got = G.unpack_pframe(bytes(b), nmb)
                G.parse_pframe(got, sc.w, sc.h)
            except G.GpuError:
                rejected += 1
    assert rejected > 0
    # streams cut short: the parse runs out of symbols and rejects
    with pytest.raises(G.GpuError):
        G.parse_pframe([s[:len(s) // 2] for s in chain[0][1]], sc.w, sc.h)


def test_parsed_codes_reconstruct_the_encoder_frame():
    """The parsed MB codes, reconstructed on the CPU from the oracle's block
    kernels (MC, intra prediction, dequantisation and inverse transforms,
    deblock), give the oracle encoder's reconstruction: the levels, flags and
    vectors the parser returns are the coded ones."""
    import codedecode as D
    for sc in (content.scene_pan(160, 96, 47, nobj=4, smin=16, smax=48),
               content.scene_blobs(128, 64, 48, nblob=30, vmax=14, grain=3.0, cut=3)):
        w, h = sc.w, sc.h
        frames = [sc.frame(i) for i in range(5)]
        near, far = frames[0], None
        for f in range(1, 5):
            navail = 1 if f == 1 else 2
            qp = (18, 26, 34, 44)[f - 1]
            st, rec, info = L.oracle_analyse(frames[f], near, far, w, h, navail, qp)
            codes = G.parse_pframe(st, w, h)
            got = D.reconstruct(codes, near, far if navail == 2 else None, w, h, qp)
            assert np.array_equal(got, rec), f"{w}x{h} frame {f}"
            far, near = near, rec
