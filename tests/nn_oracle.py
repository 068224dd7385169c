"""CPU oracle of the learned motion proposer (fdv_amd/csrc/fdvgpu_nn.hip):
forward pass, losses and parameter gradients in float64 numpy, with the
network's bf16 rounding points reproduced (weights for the forward pass, the
inputs of the three convolutions, the gradients fed back through them).  With rounding=False it is the plain float64 network, which the finite-
difference test pins.

Test infrastructure only.
"""
from __future__ import annotations

import numpy as np

G, P, C = 17, 289, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + 32 * 18
OFF_W2 = OFF_B1 + 32
OFF_B2 = OFF_W2 + 32 * 288
OFF_W3 = OFF_B2 + 32
OFF_B3 = OFF_W3 + 32 * 288
OFF_WF = OFF_B3 + 32
OFF_BF = OFF_WF + 64
NPARAM = OFF_BF + 2


def rbf(x, on=True):
    """Round to the nearest bf16 (ties to even), through float32 as the GPU does."""
    if not on:
        return np.asarray(x, dtype=np.float64)
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def _shift_index():
    """idx[tap, p] = the position p shifted by (ky-1, kx-1), or -1 (zero padding)."""
    idx = np.full((9, P), -1, dtype=np.int64)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        for p in range(P):
            y, x = p // G + dy, p % G + dx
            if 0 <= y < G and 0 <= x < G:
                idx[tap, p] = y * G + x
    return idx


SHIFT = _shift_index()


def _gather(a, tap):
    """a[shift(p, tap)] per position p (zeros outside): (P, ch)."""
    out = np.zeros_like(a)
    m = SHIFT[tap] >= 0
    out[m] = a[SHIFT[tap][m]]
    return out


def _scatter_t(a, tap):
    """The transpose of _gather: out[shift(p)] += a[p]."""
    out = np.zeros_like(a)
    m = SHIFT[tap] >= 0
    np.add.at(out, SHIFT[tap][m], a[m])
    return out


def unpack(W):
    W = np.asarray(W, dtype=np.float64)
    return dict(w1=W[OFF_W1:OFF_B1].reshape(32, 18), b1=W[OFF_B1:OFF_W2], w2=W[OFF_W2:OFF_B2].reshape(32, 288),
                b2=W[OFF_B2:OFF_W3], w3=W[OFF_W3:OFF_B3].reshape(32, 288), b3=W[OFF_B3:OFF_WF],
                wf=W[OFF_WF:OFF_BF].reshape(2, 32), bf=W[OFF_BF:NPARAM])


def _conv(x, w, b):
    """x (P, cin), w (32, 9*cin) with k = tap*cin + ci -> (P, 32)."""
    cin = x.shape[1]
    z = np.tile(b, (P, 1)).astype(np.float64)
    for tap in range(9):
        z += _gather(x, tap) @ w[:, tap * cin:(tap + 1) * cin].T
    return z


def forward(W, feat, rounding=True):
    q = unpack(W)
    r = lambda v: rbf(v, rounding)  # noqa: E731
    feat = r(np.asarray(feat, dtype=np.float64).reshape(P, 2))
    a1 = np.maximum(_conv(feat, r(q["w1"]), q["b1"]), 0.0)
    a1b = r(a1)
    a2 = np.maximum(_conv(a1b, r(q["w2"]), q["b2"]), 0.0)
    a2b = r(a2)
    a3 = np.maximum(_conv(a2b, r(q["w3"]), q["b3"]), 0.0)
    wf = r(q["wf"])
    lg = q["bf"][0] + a3 @ wf[0]
    sk = a3 @ wf[1]
    skl = sk.mean() + q["bf"][1]
    return dict(feat=feat, a1b=a1b, a2b=a2b, a3=a3, lg=lg, skl=skl, q=q, wf=wf)


def loss_and_grad(W, feats, tgts, skips, rounding=True, wfp=1.0):
    """Sums over the batch: (vector CE, SKIP CE with false positives weighed by
    wfp, gradient of their sum)."""
    r = lambda v: rbf(v, rounding)  # noqa: E731
    grad = np.zeros(NPARAM)
    lce = lsk = 0.0
    for feat, tg, y in zip(feats, tgts, skips):
        f = forward(W, feat, rounding)
        q, a3, lg = f["q"], f["a3"], f["lg"]
        m = lg.max()
        e = np.exp(lg - m)
        se = e.sum()
        ps = 1.0 / (1.0 + np.exp(-f["skl"]))
        y = float(y)
        if tg >= 0:
            lce += -(lg[tg] - m - np.log(se))
            dlg = e / se
            dlg[tg] -= 1.0
        else:
            dlg = np.zeros(P)
        lsk += -(y * np.log(max(ps, 1e-12)) + wfp * (1 - y) * np.log(max(1 - ps, 1e-12)))
        dsk = ((ps - 1.0) if y > 0.5 else wfp * ps) / P
        wf = f["wf"]
        grad[OFF_WF:OFF_WF + 32] += dlg @ a3
        grad[OFF_WF + 32:OFF_BF] += dsk * a3.sum(0)
        grad[OFF_BF] += dlg.sum()
        grad[OFF_BF + 1] += dsk * P
        dz = r((np.outer(dlg, wf[0]) + dsk * wf[1][None, :]) * (a3 > 0))
        for inp, wk, ow, ob in ((f["a2b"], q["w3"], OFF_W3, OFF_B3), (f["a1b"], q["w2"], OFF_W2, OFF_B2)):
            wr = r(wk)
            dw = np.zeros((32, 288))
            for tap in range(9):
                dw[:, tap * 32:(tap + 1) * 32] = dz.T @ _gather(inp, tap)
            grad[ow:ow + 32 * 288] += dw.ravel()
            grad[ob:ob + 32] += dz.sum(0)
            din = np.zeros((P, 32))
            for tap in range(9):
                din += _scatter_t(dz @ wr[:, tap * 32:(tap + 1) * 32], tap)
            dz = r(din * (inp > 0))
        dw1 = np.zeros((32, 18))
        for tap in range(9):
            dw1[:, tap * 2:(tap + 1) * 2] = dz.T @ _gather(f["feat"], tap)
        grad[OFF_W1:OFF_B1] += dw1.ravel()
        grad[OFF_B1:OFF_W2] += dz.sum(0)
    return lce, lsk, grad


def init_weights(seed=0, scale=1.0):
    """He-style initialisation (the training script's)."""
    rng = np.random.default_rng(seed)
    W = np.zeros(NPARAM)
    W[OFF_W1:OFF_B1] = rng.normal(0, scale * np.sqrt(2.0 / 18), 32 * 18)
    W[OFF_W2:OFF_B2] = rng.normal(0, scale * np.sqrt(2.0 / 288), 32 * 288)
    W[OFF_W3:OFF_B3] = rng.normal(0, scale * np.sqrt(2.0 / 288), 32 * 288)
    W[OFF_WF:OFF_BF] = rng.normal(0, scale * np.sqrt(1.0 / 32), 64)
    return W
