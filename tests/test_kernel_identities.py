"""Identities the HIP kernels rely on to skip work while keeping the
reference's bits, checked with IEEE double arithmetic on the host."""
import numpy as np


def test_equal_width_limiter_uses_one_division():
    # min(|RN(RN(theta dL)/d)|, |RN(RN(theta dR)/d)|) ==
    # |RN(RN(theta min(|dL|, |dR|))/d)| for d > 0 (monotone rounding)
    rng = np.random.default_rng(11)
    n = 400000
    scale = np.exp(rng.uniform(-40, 40, n))
    dL = rng.standard_normal(n) * scale
    dR = rng.standard_normal(n) * scale * np.exp(rng.uniform(-3, 3, n))
    dR[::97] = -dL[::97]            # equal magnitudes
    dR[::101] = 0.0
    dL[::103] = -0.0
    d = np.exp(rng.uniform(-30, 5, n))
    for theta in (1.4, 1.0, 2.0, -1.4):
        two = np.fmin(np.abs((theta * dL) / d), np.abs((theta * dR) / d))
        one = np.abs((theta * np.fmin(np.abs(dL), np.abs(dR))) / d)
        assert np.array_equal(two.view(np.uint64), one.view(np.uint64)), theta


def test_equal_width_limiter_special_values():
    inf, nan = np.inf, np.nan
    dL = np.array([nan, 1.0, inf, -inf, nan, 0.0])
    dR = np.array([1.0, nan, 2.0, 3.0, nan, -0.0])
    d = 0.3
    two = np.fmin(np.abs((1.4 * dL) / d), np.abs((1.4 * dR) / d))
    one = np.abs((1.4 * np.fmin(np.abs(dL), np.abs(dR))) / d)
    assert np.array_equal(two, one, equal_nan=True)
