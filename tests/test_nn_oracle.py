"""The CPU oracle of the learned proposer, pinned by what the mathematics fixes:
its gradients equal central finite differences of its own float64 loss (no
rounding), and the bf16 rounding it reproduces is exact on representable values
and ties to even."""
import numpy as np
import pytest

import nn_oracle as N


def _batch(seed, n=3):
    rng = np.random.default_rng(seed)
    feats = [np.abs(rng.normal(0, 1, (N.P, 2))) for _ in range(n)]
    tg = [int(rng.integers(0, N.P)) for _ in range(n - 1)] + [-1]
    sk = [int(rng.integers(0, 2)) for _ in range(n)]
    return feats, tg, sk


def test_bf16_rounding():
    # bf16 keeps 8 significant bits: the spacing above 1.0 is 2^-7
    x = np.array([1.0, 1.00390625, 1.01171875, 1.005859375, 1.0078125, -3.0, 0.0])
    # 1 + 2^-8 is a tie -> 1.0 (even); 1 + 3*2^-8 is a tie -> 1 + 2^-6 (even);
    # 1 + 3*2^-9 rounds up to 1 + 2^-7; 1 + 2^-7 is representable
    assert np.array_equal(N.rbf(x), [1.0, 1.0, 1.015625, 1.0078125, 1.0078125, -3.0, 0.0])


@pytest.mark.parametrize("seed", [1, 2])
def test_gradients_match_finite_differences(seed):
    W = N.init_weights(seed, scale=1.0)
    W[N.OFF_B1:N.OFF_W2] = 0.1          # keep ReLUs away from their kinks
    W[N.OFF_B2:N.OFF_W3] = 0.1
    W[N.OFF_B3:N.OFF_WF] = 0.1
    feats, tg, sk = _batch(seed)
    lce, lsk, g = N.loss_and_grad(W, feats, tg, sk, rounding=False, wfp=2.0)
    rng = np.random.default_rng(100 + seed)
    idx = np.concatenate([rng.integers(s, e, 6) for s, e in
                          ((N.OFF_W1, N.OFF_B1), (N.OFF_B1, N.OFF_W2), (N.OFF_W2, N.OFF_B2), (N.OFF_W3, N.OFF_B3),
                           (N.OFF_B3, N.OFF_WF), (N.OFF_WF, N.OFF_BF))] + [np.array([N.OFF_BF, N.OFF_BF + 1])])
    h = 1e-5
    for k in idx:
        Wp, Wm = W.copy(), W.copy()
        Wp[k] += h
        Wm[k] -= h
        lp = sum(N.loss_and_grad(Wp, feats, tg, sk, rounding=False, wfp=2.0)[:2])
        lm = sum(N.loss_and_grad(Wm, feats, tg, sk, rounding=False, wfp=2.0)[:2])
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 + 1e-4 * abs(fd), (int(k), fd, g[k])
