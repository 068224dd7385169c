"""The BD-rate method (tools/rd-bench.py:75-113 restated in fdv_amd/learned/
bdrate.py), pinned by closed forms: identical curves give 0 %, scaling every
rate by k gives (k - 1) * 100 %, dominated points are dropped."""
import numpy as np
import pytest

from fdv_amd.learned import bdrate as B

CURVE = [(1000.0, 30.0), (1800.0, 33.0), (3400.0, 36.2), (6500.0, 39.5)]


def test_identical_curves():
    assert abs(B.bd_rate(CURVE, CURVE)) < 1e-9


@pytest.mark.parametrize("k", [0.8, 1.015, 1.5])
def test_scaled_rates(k):
    assert B.bd_rate(CURVE, [(r * k, d) for r, d in CURVE]) == pytest.approx((k - 1) * 100, abs=1e-6)


def test_dominated_points_dropped():
    pts = CURVE + [(2500.0, 32.0)]            # costs more than (1800, 33) for less PSNR
    assert B.frontier(pts) == sorted(CURVE)
    assert abs(B.bd_rate(CURVE, pts)) < 1e-9


def test_too_few_points():
    assert B.bd_rate(CURVE[:2], CURVE) is None


def test_shifted_quality():
    # a new curve 0.5 dB better everywhere: negative BD-rate
    assert B.bd_rate(CURVE, [(r, d + 0.5) for r, d in CURVE]) < 0


def test_psnr():
    a = np.zeros(100, np.uint8)
    assert B.psnr_y(a, a, 100) == 99.0
    assert B.psnr_y(a + 1, a, 100) == pytest.approx(10 * np.log10(255.0 ** 2))
