"""The block-parallel modes at the measured workloads' real shapes, on the GPU
against the oracle, with every payload decoded by the reference decoder:

- the learned proposer on the 1280x720 eight-band workload (pan over a
  blurred-noise texture, 12 rectangles, grain sigma 3, QP 24);
- sixteen independent streams in one batch: 8 x 640x368, 4 x 1280x720,
  4 x 1920x1088, QP per stream from {20, 24, 28, 32};
- the stress content (500 fast blobs, grain sigma 5) at 1920x1088 in 6-MB-row
  bands, learned mode.

Per sequence and frame: the GPU's motion table equals the oracle's candidate
search given the GPU's proposals; streams, reconstruction and decisions equal
the oracle's block-parallel coding of that table; and the payload packed from
the GPU's streams (entropy state carried) decodes with the reference decoder
(pframe_decode, fdv.h:6492) to exactly the GPU's reconstruction.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G
from gpucases import mbinfo_to_np, same_decisions, say

pytestmark = pytest.mark.gpu


def bands(h, band_mbrows=6):
    out, y = [], 0
    while y < h:
        out.append((y, min(band_mbrows * 16, h - y)))
        y += band_mbrows * 16
    return out


def band_frames(frame, w, h, layout):
    ys, cw, ch = w * h, w // 2, h // 2
    Y = frame[:ys].reshape(h, w)
    U = frame[ys:ys + cw * ch].reshape(ch, cw)
    V = frame[ys + cw * ch:].reshape(ch, cw)
    return [np.concatenate([Y[y0:y0 + bh].ravel(), U[y0 // 2:(y0 + bh) // 2].ravel(),
                            V[y0 // 2:(y0 + bh) // 2].ravel()]) for y0, bh in layout]


def run_learned(sizes, seq_frames, qps, label, decode=True):
    """Learned-mode frames of every sequence in one batch, checked against the
    oracle and, with `decode`, through the reference decoder.  Returns the
    number of MBs that took the candidate search."""
    assert L.ref() is not None or not decode, "the reference build (oracle/_ref) is needed for the decode check"
    n = len(sizes)
    nfr = len(seq_frames[0])
    an = G.Analyser(sizes, nsrc=1)
    an.set_proposer()
    near = [seq_frames[k][0] for k in range(n)]
    far = [None] * n
    tce = [L.TabCache() for _ in range(n)]
    tcd = [L.TabCache() for _ in range(n)]
    for k in range(n):
        an.put_ref(k, 0, near[k])
    ncand = 0
    with ThreadPoolExecutor(16) as ex:
        for f in range(1, nfr):
            navail = [1 if f == 1 else 2] * n
            for k in range(n):
                an.put_source(k, 0, seq_frames[k][f])
            say(f"  [{label}] frame {f}: GPU learned step over {n} sequences")
            an.pframe([0] * n, qps, navail=navail, me_mode=G.ME_LEARNED)
            an.fetch()
            got = [(an.proposals(k), an.motion(k), an.streams(k), an.recon(k), an.decisions(k)) for k in range(n)]

            def check(k):
                w, h = sizes[k]
                props, gm, gst, grec, gdec = got[k]
                cur = seq_frames[k][f]
                mt = L.oracle_metab_candidates(cur, near[k], far[k], w, h, navail[k], props)
                om = L.metab_to_np(mt)
                bad = [i for i in range(len(om)) if om[i].tobytes() != gm[i].tobytes()]
                if bad:
                    return f"seq {k} {w}x{h}: motion table differs at {len(bad)} MBs, first {bad[0]}"
                st, rec, info = L.oracle_parallel(cur, near[k], far[k], w, h, navail[k], qps[k], metab=mt)
                if gst != st:
                    return f"seq {k} {w}x{h}: streams differ {[j for j in range(16) if gst[j] != st[j]]}"
                if not np.array_equal(grec, rec):
                    return f"seq {k} {w}x{h}: reconstruction differs"
                di = same_decisions(gdec, mbinfo_to_np(info), (w // 16) * (h // 16))
                if di >= 0:
                    return f"seq {k} {w}x{h}: decision of MB {di} differs"
                if decode:
                    pay = L.oracle_pack(qps[k], gst, tce[k])
                    dec = L.ref_pframe_decode(pay, near[k], far[k], w, h, navail[k], tcd[k])
                    if dec is None:
                        return f"seq {k} {w}x{h}: the reference decoder rejected the payload"
                    if not np.array_equal(dec, grec):
                        return f"seq {k} {w}x{h}: reference decoder output differs from the GPU reconstruction"
                return None

            errs = [e for e in ex.map(check, range(n)) if e]
            assert not errs, f"[{label}] frame {f}: " + "; ".join(errs[:4])
            ncand += sum(int((g[0]["conf"] >= 0.7).sum()) for g in got)
            for k in range(n):
                far[k], near[k] = near[k], got[k][3]
    an.close()
    return ncand


def test_learned_720p_eight_bands():
    W, H = 1280, 720
    layout = bands(H)
    sc = content.scene_pan(W, H, 2)
    per = [band_frames(sc.frame(i), W, H, layout) for i in range(4)]
    seq_frames = [[per[f][b] for f in range(4)] for b in range(len(layout))]
    ncand = run_learned([(W, bh) for _, bh in layout], seq_frames, [24] * len(layout), "720p-learned")
    assert ncand > 0


def test_batched_sixteen_streams_real_sizes():
    rng = np.random.default_rng(2024)
    sizes = [(640, 368)] * 8 + [(1280, 720)] * 4 + [(1920, 1088)] * 4
    qps = [int(q) for q in rng.choice([20, 24, 28, 32], size=len(sizes))]
    seq_frames = []
    for k, (w, h) in enumerate(sizes):
        sc = content.scene_objects(w, h, 300 + k, nobj=4 + w // 320, smin=24, smax=96, vmax=8, grain=1.0)
        seq_frames.append([sc.frame(i) for i in range(3)])
    assert run_learned(sizes, seq_frames, qps, "16-streams") > 0


def test_stress_1088p_learned_bands():
    W, H = 1920, 1088
    layout = bands(H)
    sc = content.scene_blobs(W, H, 5)
    per = [band_frames(sc.frame(i), W, H, layout) for i in range(3)]
    seq_frames = [[per[f][b] for f in range(3)] for b in range(len(layout))]
    run_learned([(W, bh) for _, bh in layout], seq_frames, [24] * len(layout), "stress-1088p")
