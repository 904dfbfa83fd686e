"""Pins of the oracle's weighted BasicStatistics branch (refcpu::weighted_statistics,
BasicStatistics.cc:144-188, resample :50-73, weighted_variance misc/math.hh:45-54).

The reference's tests of this branch only construct an empty weight sampler
(python/tests/sampling/test_CompletionCheck.py); the restatement is pinned
here by closed forms:
  * resample: eq[i] is the observation whose cumulative weight interval holds
    i * W / n, with the reference's strict `<` (an index i equal to a
    cumulative weight stays on the earlier observation), checked on dyadic
    weights where every sum is exact;
  * method 1: mean = sum(x w) / W and variance = sum(w (x - mean)^2) / W;
    with unit weights both equal the unweighted ones;
  * the autocorrelation factor of method 1 is that of the resampled series
    with rho = 2^(-1 / (lag * W / n)), checked on weights whose resampled
    series is a square wave with a hand-counted lag;
  * method 2 is the unweighted statistics of the resampled series;
  * time-grid records (weight = number of grid points) average like the
    expanded grid.
"""
import numpy as np
import pytest

import oracle


def _ar1(n, seed, phi=0.9):
    rng = np.random.default_rng(seed)
    x = np.empty(n)
    x[0] = rng.normal()
    for i in range(1, n):
        x[i] = phi * x[i - 1] + rng.normal()
    return x + 5.0


def test_resample_closed_form_dyadic():
    rng = np.random.default_rng(3)
    x = rng.normal(size=200)
    w = rng.integers(1, 9, size=200) * 0.25  # dyadic: every partial sum is exact
    W = float(w.sum())
    for n in (int(W * 4), int(W), 137):
        eq = oracle.resample(x, w, W, n)
        C = np.cumsum(w)
        targets = np.arange(n) * (W / n)
        j = np.searchsorted(C, targets, side="left")  # first j with C_j >= target
        np.testing.assert_array_equal(eq, x[j])


def test_resample_boundary_stays_on_earlier_observation():
    x = np.array([10.0, 20.0, 30.0])
    w = np.array([2.0, 1.0, 1.0])
    # W = 4, n = 4: targets 0, 1, 2, 3 -> 10, 10, 10 (2 is the end of the first
    # interval), 20
    np.testing.assert_array_equal(oracle.resample(x, w, 4.0, 4), [10.0, 10.0, 10.0, 20.0])


@pytest.mark.parametrize("seed", [1, 2])
def test_method1_mean_and_variance(seed):
    x = _ar1(3000, seed)
    w = np.random.default_rng(seed + 10).exponential(size=3000)
    s = oracle.weighted_statistics(x, w, 0.95, 1, 10000)
    W = w.sum()
    m = (x * w).sum() / W
    assert s["mean"] == pytest.approx(m, rel=1e-13)
    assert s["variance"] == pytest.approx((w * (x - m) ** 2).sum() / W, rel=1e-12)
    assert s["calculated_precision"] == pytest.approx(
        np.sqrt(2.0) * _erfinv(0.95) * np.sqrt(s["autocorr_factor"] * s["variance"] / W), rel=1e-12)


def _erfinv(x):  # math.hh:62-72, restated for the precision check
    a = 0.147
    b = np.log((1 - x) * (1 + x))
    c = 2 / (np.pi * a) + b / 2
    return np.sign(x) * np.sqrt(np.sqrt(c * c - b / a) - c)


def test_unit_weights_reduce_to_unweighted():
    x = _ar1(2000, 7)
    s = oracle.weighted_statistics(x, np.ones_like(x), 0.95, 1, 2000)
    assert s["mean"] == pytest.approx(x.mean(), rel=1e-13)
    assert s["variance"] == pytest.approx(x.var(), rel=1e-12)


def test_increment_scales_the_lag():
    """Observations 11, 9 with weights 1, 1.5 (W = 2.5) read at 10 points i * 0.25
    (all exact): points 0..4 lie in the first interval [0, 1] (the strict < keeps
    i = 4, at 1.0, on it) and 5..9 in the second, so the resampled series is
    10 + (+1 x 5, -1 x 5), whose first lag with |cov/var0| <= 0.5 is 2
    (cov_1 = 7/9, cov_2 = 4/8).  With increment W/n = 0.25, rho = 2^(-1/(2*0.25))
    = 1/4 and the factor is (1 + 1/4)/(1 - 1/4) = 5/3.  Weighted mean
    (11 + 13.5)/2.5 = 9.8, variance (1.2^2 + 1.5 * 0.8^2)/2.5 = 0.96."""
    x = np.array([11.0, 9.0])
    w = np.array([1.0, 1.5])
    eq = oracle.resample(x, w, 2.5, 10)
    assert list(eq) == [11.0] * 5 + [9.0] * 5
    s = oracle.weighted_statistics(x, w, 0.95, 1, 10)
    assert s["autocorr_factor"] == pytest.approx(5.0 / 3.0, rel=1e-14)
    assert s["mean"] == pytest.approx(9.8, rel=1e-15)
    assert s["variance"] == pytest.approx(0.96, rel=1e-14)
    fc = np.sqrt(2.0) * oracle.lib().orc_approx_erf_inv(0.95)
    assert s["calculated_precision"] == pytest.approx(fc * np.sqrt(5.0 / 3.0 * 0.96 / 2.5),
                                                      rel=1e-14)


def test_method2_is_unweighted_statistics_of_resampled():
    x = _ar1(1200, 11)
    w = np.random.default_rng(12).exponential(size=1200)
    s = oracle.weighted_statistics(x, w, 0.9, 2, 5000)
    eq = oracle.resample(x, w, w.sum(), 5000)
    m, p = oracle.basic_statistics(eq, 0.9)
    assert s["mean"] == pytest.approx(m, rel=1e-14)
    assert s["calculated_precision"] == pytest.approx(p, rel=1e-12)


def test_time_grid_records_average_like_the_expanded_grid():
    rng = np.random.default_rng(13)
    x = rng.integers(-50, 50, size=400).astype(float)
    w = rng.integers(1, 6, size=400).astype(float)
    s = oracle.weighted_statistics(x, w, 0.95, 1, 10000)
    assert s["mean"] == pytest.approx(np.repeat(x, w.astype(int)).mean(), rel=1e-13)


def test_errors():
    with pytest.raises(ValueError):
        oracle.weighted_statistics([], [], 0.95, 1, 100)
    with pytest.raises(ValueError):
        oracle.weighted_statistics([1.0, 2.0], [1.0], 0.95, 1, 100)
    with pytest.raises(ValueError):
        oracle.weighted_statistics([1.0, 2.0], [1.0, 1.0], 0.95, 3, 100)
