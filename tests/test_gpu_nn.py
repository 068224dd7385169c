"""The learned proposer on the GPU: its training kernels against the float64
CPU oracle (tests/nn_oracle.py, itself pinned by finite differences), its
inference against the oracle's forward pass, and the learned encode mode bit-
exact with the oracle given the same proposals."""
import ctypes as C

import numpy as np
import pytest

import nn_oracle as N
import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G
from gpucases import mbinfo_to_np, same_decisions, say

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _weights(seed):
    W = N.init_weights(seed, scale=1.0)
    rng = np.random.default_rng(seed)
    W[N.OFF_B1:N.OFF_W2] = rng.normal(0, 0.1, 32)
    W[N.OFF_B2:N.OFF_W3] = rng.normal(0, 0.1, 32)
    W[N.OFF_B3:N.OFF_WF] = rng.normal(0, 0.1, 32)
    W[N.OFF_BF:N.NPARAM] = [0.1, -0.2]
    return W.astype(np.float32)


def test_train_step_gradients_match_oracle():
    """Loss and every parameter group's gradient: relative L2 error <= 1e-2
    (bf16 rounding points are the same on both sides; the rest is fp32 vs
    float64 accumulation and the occasional bf16 rounding flip)."""
    lib = G.lib()
    rng = np.random.default_rng(5)
    n = 6
    feats = np.abs(rng.normal(0, 1.5, (n, N.P, 2))).astype(np.float32)
    tg = np.array([int(rng.integers(0, N.P)) for _ in range(n - 1)] + [-1], dtype=np.int32)
    sk = np.array([1, 0, 1, 0, 0, 1], dtype=np.uint8)
    W = _weights(3)
    dW = torch.tensor(W, device="cuda")
    dF = torch.tensor(feats.reshape(n, -1), device="cuda")
    dT = torch.tensor(tg, device="cuda")
    dS = torch.tensor(sk, device="cuda")
    grad = torch.zeros(N.NPARAM, dtype=torch.float32, device="cuda")
    loss = torch.zeros(8, dtype=torch.float32, device="cuda")
    say("  nn train step: 6 tuples, 3 workgroups")
    assert lib.fdvgpu_nn_train_step(dW.data_ptr(), dF.data_ptr(), dT.data_ptr(), dS.data_ptr(), n, C.c_float(2.5),
                                    C.c_float(0.7), grad.data_ptr(), loss.data_ptr(), 3, None) == 0
    torch.cuda.synchronize()
    g = grad.cpu().numpy().astype(np.float64)
    lgpu = loss.cpu().numpy()
    lce, lsk, go = N.loss_and_grad(W.astype(np.float64), list(feats), list(tg), list(sk), rounding=True, wfp=2.5)
    assert lgpu[0] == pytest.approx(lce, rel=1e-3)
    assert lgpu[1] == pytest.approx(lsk, rel=1e-3)
    groups = {"w1": (N.OFF_W1, N.OFF_B1), "b1": (N.OFF_B1, N.OFF_W2), "w2": (N.OFF_W2, N.OFF_B2),
              "b2": (N.OFF_B2, N.OFF_W3), "w3": (N.OFF_W3, N.OFF_B3), "b3": (N.OFF_B3, N.OFF_WF),
              "wf": (N.OFF_WF, N.OFF_BF), "bf": (N.OFF_BF, N.NPARAM)}
    for name, (a, b) in groups.items():
        err = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-12)
        assert err <= 1e-2, (name, err)


def test_adam_matches_formula():
    lib = G.lib()
    rng = np.random.default_rng(9)
    n = 1000
    W = rng.normal(0, 1, n).astype(np.float32)
    gr = rng.normal(0, 1, n).astype(np.float32)
    m1 = rng.normal(0, 0.1, n).astype(np.float32)
    m2 = np.abs(rng.normal(0, 0.1, n)).astype(np.float32)
    t = [torch.tensor(x, device="cuda") for x in (W, gr, m1, m2)]
    assert lib.fdvgpu_nn_adam(t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), t[3].data_ptr(), n,
                              C.c_float(1e-3), 3, 4, None) == 0
    torch.cuda.synchronize()
    g = gr.astype(np.float64) / 4
    a = 0.9 * m1 + 0.1 * g
    b = 0.999 * m2 + 0.001 * g * g
    want = W - 1e-3 * (a / (1 - 0.9 ** 3)) / (np.sqrt(b / (1 - 0.999 ** 3)) + 1e-8)
    assert np.allclose(t[0].cpu().numpy(), want, rtol=1e-5, atol=1e-7)
    assert not t[1].cpu().numpy().any()


def _gpu_features(an, seq, w, h, qp):
    lib = G.lib()
    nmb = (w // 16) * (h // 16)
    cy, _, ry, rs = an.planes(seq, 0)
    pos = torch.tensor([[(i % (w // 16)) * 16, (i // (w // 16)) * 16] for i in range(nmb)], dtype=torch.int32,
                       device="cuda")
    curs = torch.full((nmb,), cy, dtype=torch.int64, device="cuda")
    refs = torch.full((nmb,), ry, dtype=torch.int64, device="cuda")
    cw = torch.full((nmb,), w, dtype=torch.int32, device="cuda")
    rss = torch.full((nmb,), rs, dtype=torch.int32, device="cuda")
    qps = torch.full((nmb,), qp, dtype=torch.int32, device="cuda")
    feats = torch.zeros((nmb, N.P * 2), dtype=torch.float32, device="cuda")
    assert lib.fdvgpu_nn_features(curs.data_ptr(), cw.data_ptr(), refs.data_ptr(), rss.data_ptr(), pos.data_ptr(),
                                  qps.data_ptr(), nmb, feats.data_ptr(), None) == 0
    torch.cuda.synchronize()
    return feats.cpu().numpy().reshape(nmb, N.P, 2)


def _quadrant_lowest_restated(cur, ref, w, h):
    """Per MB and 8x8 quadrant, the 2 lowest-cost positions of the quadrant's own
    half-res correlation map (ties: |dx| + |dy|, then position), in pels."""
    Y = cur[:w * h].reshape(h, w).astype(np.int64)
    R = np.pad(ref[:w * h].reshape(h, w), 32, mode="edge").astype(np.int64)
    out = []
    for i in range((w // 16) * (h // 16)):
        bx, by = (i % (w // 16)) * 16, (i // (w // 16)) * 16
        b = Y[by:by + 16, bx:bx + 16]
        c8 = (b[0::2, 0::2] + b[0::2, 1::2] + b[1::2, 0::2] + b[1::2, 1::2] + 2) >> 2
        win = R[32 + by - 24:32 + by + 40, 32 + bx - 24:32 + bx + 40]
        r32 = (win[0::2, 0::2] + win[0::2, 1::2] + win[1::2, 0::2] + win[1::2, 1::2] + 2) >> 2
        mb = []
        for q in range(4):
            oy, ox = (q >> 1) * 4, (q & 1) * 4
            keys = []
            for p in range(289):
                dy, dx = p // 17 - 8, p % 17 - 8
                c = np.abs(c8[oy:oy + 4, ox:ox + 4] - r32[12 + dy + oy:16 + dy + oy, 12 + dx + ox:16 + dx + ox]).sum()
                keys.append((int(c), abs(dx) + abs(dy), p))
            mb.append([(2 * (k[2] % 17 - 8), 2 * (k[2] // 17 - 8)) for k in sorted(keys)[:2]])
        out.append(mb)
    return out


def _features_restated(cur, ref, w, h, qp):
    """The correlation map of every MB in numpy (the definition in fdvgpu_nn.hip)."""
    Y = cur[:w * h].reshape(h, w).astype(np.int64)
    R = np.pad(ref[:w * h].reshape(h, w), 32, mode="edge").astype(np.int64)
    out = []
    for i in range((w // 16) * (h // 16)):
        bx, by = (i % (w // 16)) * 16, (i // (w // 16)) * 16
        b = Y[by:by + 16, bx:bx + 16]
        c8 = (b[0::2, 0::2] + b[0::2, 1::2] + b[1::2, 0::2] + b[1::2, 1::2] + 2) >> 2
        win = R[32 + by - 24:32 + by + 40, 32 + bx - 24:32 + bx + 40]
        r32 = (win[0::2, 0::2] + win[0::2, 1::2] + win[1::2, 0::2] + win[1::2, 1::2] + 2) >> 2
        cost = np.array([[np.abs(c8 - r32[12 + dy:20 + dy, 12 + dx:20 + dx]).sum() for dx in range(-8, 9)]
                         for dy in range(-8, 9)]).ravel()
        s0 = np.float32(1.0) / (np.float32(16.0) * np.float32(2.0 ** ((qp - 4) / 6.0)))
        out.append(np.stack([cost * s0, (cost - cost.min()) / 256.0], 1))
    return np.array(out)


def test_features_and_inference_match_oracle():
    sc = content.scene_pan(96, 64, 41, nobj=3, smin=8, smax=32)
    w, h = sc.w, sc.h
    f0, f1 = sc.frame(0), sc.frame(1)
    an = G.Analyser([(w, h)], nsrc=1)
    W = _weights(11)
    an.set_proposer(W)
    an.put_ref(0, 0, f0)
    an.put_source(0, 0, f1)
    feats = _gpu_features(an, 0, w, h, 26)
    want = _features_restated(f1, f0, w, h, 26)
    quads = _quadrant_lowest_restated(f1, f0, w, h)
    assert np.allclose(feats, want, rtol=1e-6, atol=1e-6)
    say("  nn inference: learned frame 96x64")
    an.pframe([0], [26], navail=[1], me_mode=G.ME_LEARNED)
    an.fetch()
    props = an.proposals(0)
    for i in range(len(props)):
        fw = N.forward(W.astype(np.float64), feats[i], rounding=True)
        lg = fw["lg"]
        order = np.argsort(-lg, kind="stable")
        top = [(int(2 * (p % 17 - 8)), int(2 * (p // 17 - 8))) for p in order[:8]]
        got = list(zip(props[i]["cx"].tolist(), props[i]["cy"].tolist()))
        # the leading vectors agree (later ranks may swap on near-ties)
        assert got[0] == top[0] or abs(lg[order[0]] - lg[order[1]]) < 1e-3
        e = np.exp(lg - lg.max())
        assert props[i]["conf"] == pytest.approx(e[order[:8]].sum() / e.sum(), abs=2e-3)
        assert props[i]["pskip"] == pytest.approx(1 / (1 + np.exp(-fw["skl"])), abs=2e-3)
        # the correlation map's 4 lowest costs: ties to the smaller |dx| + |dy|, then the lower position
        rel = np.rint(want[i][:, 1] * 256).astype(np.int64)
        pos = np.arange(289)
        key = [(int(rel[p]), abs(p % 17 - 8) + abs(p // 17 - 8), int(p)) for p in pos]
        low = [k[2] for k in sorted(key)[:4]]
        assert list(zip(props[i]["kx"].tolist(), props[i]["ky"].tolist())) == \
            [(2 * (p % 17 - 8), 2 * (p // 17 - 8)) for p in low]
        for q in range(4):
            assert list(zip(props[i]["qx"][q].tolist(), props[i]["qy"][q].tolist())) == quads[i][q], (i, q)
    an.close()


@pytest.mark.parametrize("weights_seed", [1, 2])
def test_learned_mode_bit_exact_with_oracle(weights_seed):
    """Given the GPU's proposals, the oracle's candidate search, decisions and
    coding reproduce the GPU's tables, streams, reconstruction and decisions."""
    scenes = [content.scene_objects(96, 64, 51, nobj=3, smin=8, smax=32, vmax=6),
              content.scene_pan(64, 48, 52, nobj=2, smin=8, smax=24, pan=(2, -1))]
    sizes = [(s.w, s.h) for s in scenes]
    an = G.Analyser(sizes, nsrc=1)
    W = _weights(weights_seed)
    W[N.OFF_WF:N.OFF_BF] *= 4.0                      # sharper logits: both search branches occur
    an.set_proposer(W)
    sthr = 0.7 if weights_seed == 1 else 2.0         # SKIP pruning on, then off (the default)
    if weights_seed == 1:
        an.set_proposer_thresholds(0.7, sthr)
    frames = [[s.frame(i) for i in range(4)] for s in scenes]
    near = [fr[0] for fr in frames]
    far = [None, None]
    for k in range(2):
        an.put_ref(k, 0, near[k])
    ncand = nfull = 0
    for f in range(1, 4):
        navail = [1 if f == 1 else 2] * 2
        for k in range(2):
            an.put_source(k, 0, frames[k][f])
        say(f"  learned mode frame {f}")
        an.pframe([0, 0], [24, 30], navail=navail, me_mode=G.ME_LEARNED)
        an.fetch()
        for k, (w, h) in enumerate(sizes):
            props = an.proposals(k)
            ncand += int((props["conf"] >= 0.7).sum())
            nfull += int((props["conf"] < 0.7).sum())
            mt = L.oracle_metab_candidates(frames[k][f], near[k], far[k], w, h, navail[k], props, sthr=sthr)
            om, gm = L.metab_to_np(mt), an.motion(k)
            bad = [i for i in range(len(om)) if om[i].tobytes() != gm[i].tobytes()]
            assert not bad, (f"seq {k} frame {f}: motion table differs at {len(bad)} MBs, first {bad[0]}: "
                             f"oracle {om[bad[0]]} gpu {gm[bad[0]]} proposal {props[bad[0]]}")
            st, rec, info = L.oracle_parallel(frames[k][f], near[k], far[k], w, h, navail[k], [24, 30][k],
                                                    metab=mt)
            assert an.streams(k) == st, f"seq {k} frame {f}: streams"
            assert np.array_equal(an.recon(k), rec), f"seq {k} frame {f}: recon"
            assert same_decisions(an.decisions(k), mbinfo_to_np(info), (w // 16) * (h // 16)) < 0
            far[k], near[k] = near[k], rec
    an.close()
    assert ncand > 0


def test_thresholds_above_one_give_the_exhaustive_table():
    """With both proposer thresholds above 1 no MB takes the candidate search or
    the SKIP pruning: the learned mode's motion table is the exhaustive one."""
    sc = content.scene_objects(96, 64, 61, nobj=3, smin=8, smax=32, vmax=6)
    w, h = sc.w, sc.h
    an = G.Analyser([(w, h)], nsrc=2)
    an.set_proposer(_weights(3))
    an.set_proposer_thresholds(2.0, 2.0)
    f0, f1, f2 = sc.frame(0), sc.frame(1), sc.frame(2)
    an.put_ref(0, 0, f0)
    an.put_source(0, 0, f1)
    an.put_source(0, 1, f2)
    say("  thresholds above 1: frame 1")
    an.pframe([0], [24], navail=[1], me_mode=G.ME_LEARNED)
    an.fetch()
    rec1 = an.recon(0)
    om = L.metab_to_np(L.oracle_metab_exhaustive(f1, f0, None, w, h, 1))
    gm = an.motion(0)
    assert all(om[i].tobytes() == gm[i].tobytes() for i in range(len(om)))
    say("  thresholds above 1: frame 2")
    an.pframe([1], [24], navail=[2], me_mode=G.ME_LEARNED)
    an.fetch()
    om = L.metab_to_np(L.oracle_metab_exhaustive(f2, rec1, f0, w, h, 2))
    gm = an.motion(0)
    assert all(om[i].tobytes() == gm[i].tobytes() for i in range(len(om)))
    an.close()
