"""Pin the oracle's statistics and equilibration restatements by results the
mathematics fixes (CPU only).

Functions pinned here (oracle/refcpu.hh):
  * autocorrelation_factor (BasicStatistics.cc:24-48): deterministic square
    waves whose lag covariances are counted by hand, so the first lag with
    |cov_k / var0| <= 0.5 and the factor (1 + rho) / (1 - rho), rho = 2^(-1/k),
    are known in closed form (k = 1 gives 3, k = 2 gives 3 + 2 sqrt 2); the
    |var0 / mean| < 1e-8 early-out; the DBL_MAX result when no lag qualifies;
    the increment of resampled weighted observations.
  * default_equilibration_check_impl (EquilibrationCheck.cc:50-117): series
    whose partition scan and mean-crossing step are worked out by hand below,
    chosen so that each plausible misreading (the second partition starting at
    N/2 for odd N, ">=" for ">", the wrong parity of the first move, the mean of
    the whole series in the crossing step, a missing or reversed crossing step,
    "start1 < N" as the equilibrated test, an absolute eps in the all-same
    test) changes at least one expected result.
  * basic_statistics / approx_erf_inv (BasicStatistics.cc:114-131,
    misc/math.hh:62-72): the calculated precision of a series with a known
    factor, and Winitzki's approximation against scipy's erfinv within the
    approximation's own error bound.
  * CompletionCheck with requested precisions: the behavioural assertions of
    the reference's python/tests/sampling/test_CompletionCheck.py:40-124,
    replayed on seeded engines.

Square waves used for the lag search: d_j = +1 for j mod 2p < p, else -1, over
N = 2p (one period).  For lag i <= p the products d_j d_{j+i}, j < N - i, are
-1 exactly for p - i <= j < p (i of them) and +1 for the other N - 2i, so
cov_i = (N - 3i) / (N - i) with var0 = 1 (mean of the +-1 part is 0).
"""
import math

import numpy as np
import pytest

import oracle as O

DBL_MAX = np.finfo(np.float64).max


def square(p, periods=1):
    return np.array(([1.0] * p + [-1.0] * p) * periods)


def factor(k, increment=1.0):
    rho = 2.0 ** (-1.0 / (k * increment))
    return (1.0 + rho) / (1.0 - rho)


# --- autocorrelation_factor --------------------------------------------------------

def test_lag_at_exact_half_is_taken():
    """p = 5, N = 10: cov_1 = 7/9, cov_2 = (10 - 6)/8 = 0.5 exactly, and the
    reference's test is <= 0.5, so the lag is 2 and the factor is
    (1 + 2^-1/2) / (1 - 2^-1/2) = 3 + 2 sqrt 2.  (A strict < would give lag 3.)"""
    x = 10.0 + square(5)
    assert O.autocorrelation_factor(x) == pytest.approx(3.0 + 2.0 * math.sqrt(2.0), rel=1e-14)


def test_covariance_is_normalised_by_the_overlap():
    """p = 8, N = 16: cov_3 = 7/13 > 0.5 and cov_4 = 4/12 <= 0.5, so the lag is 4.
    Dividing the lag sums by N instead of N - i would give 7/16 at lag 3."""
    x = 10.0 + square(8)
    assert O.autocorrelation_factor(x) == pytest.approx(factor(4), rel=1e-14)
    assert O.autocorrelation_factor(x) != pytest.approx(factor(3), rel=1e-6)


def test_lag_one_gives_three():
    """d = (+1, +1, -1, -1) x 3, N = 12: adjacent products are +1 at even j (6) and
    -1 at odd j (5), cov_1 = 1/11, so the lag is 1, rho = 1/2 and the factor is 3."""
    x = 10.0 + np.array([1.0, 1.0, -1.0, -1.0] * 3)
    assert O.autocorrelation_factor(x) == 3.0


def test_no_qualifying_lag_gives_dbl_max():
    """An alternating series has cov_i = (-1)^i at every lag: no |cov| <= 0.5."""
    x = 3.0 + np.array([1.0, -1.0] * 4)
    assert O.autocorrelation_factor(x) == DBL_MAX


def test_negative_mean_uses_the_absolute_ratio():
    """|var0 / mean| = 0.1 for mean -10: no early-out (a signed ratio would be
    below 1e-8 and return 1)."""
    x = -10.0 + square(8)
    assert O.autocorrelation_factor(x) == pytest.approx(factor(4), rel=1e-14)


def test_zero_mean_is_not_an_early_out():
    """mean 0: var0 / mean is +inf, not < 1e-8, so the lag search runs."""
    assert O.autocorrelation_factor(square(8)) == pytest.approx(factor(4), rel=1e-14)


def test_no_variation_early_out():
    # var0 == 0
    assert O.autocorrelation_factor(np.full(20, 5.0)) == 1.0
    # var0 / mean ~ 1e-4 / 1e6 = 1e-10 < 1e-8
    assert O.autocorrelation_factor(1.0e6 + 1.0e-2 * square(8)) == 1.0
    # var0 / mean ~ 1e-4 / 1e3 = 1e-7: the lag search runs (lag 4)
    assert O.autocorrelation_factor(1.0e3 + 1.0e-2 * square(8)) == pytest.approx(
        factor(4), rel=1e-12)


@pytest.mark.parametrize("increment,expected", [(0.5, 3.0), (2.0, 3.0 + 2.0 * math.sqrt(2.0))])
def test_increment_of_resampled_observations(increment, expected):
    """rho = 2^(-1/(lag * increment)): lag 2 at increment 1/2 gives rho = 1/2
    (factor 3); lag 1 at increment 2 gives rho = 2^-1/2 (3 + 2 sqrt 2)."""
    x = 10.0 + (square(5) if increment == 0.5 else np.array([1.0, 1.0, -1.0, -1.0] * 3))
    assert O.autocorrelation_factor(x, increment) == pytest.approx(expected, rel=1e-14)


# --- basic_statistics and approx_erf_inv --------------------------------------------

def test_erf_inv_approximation_within_its_bound():
    """Winitzki's a = 0.147 approximation of erf^-1 stays within 2e-3 relative of
    the exact inverse (scipy) on (0, 1); at 0.95 sqrt(2) erf^-1 is 1.95996."""
    from scipy.special import erfinv
    for x in np.linspace(0.01, 0.99, 99):
        assert abs(O.lib().orc_approx_erf_inv(float(x)) / erfinv(x) - 1.0) < 2e-3
    assert O.lib().orc_approx_erf_inv(-0.5) == -O.lib().orc_approx_erf_inv(0.5)
    assert math.sqrt(2.0) * O.lib().orc_approx_erf_inv(0.95) == pytest.approx(1.95996, rel=1e-3)


def test_calculated_precision_of_a_known_factor():
    """Series of test_lag_one_gives_three: mean 10, var0 1, factor 3, N 12, so
    the precision is sqrt(2) erf^-1(c) * sqrt(3 / 12) = half the confidence factor."""
    x = 10.0 + np.array([1.0, 1.0, -1.0, -1.0] * 3)
    m, p = O.basic_statistics(x, 0.95)
    assert m == 10.0
    fc = math.sqrt(2.0) * O.lib().orc_approx_erf_inv(0.95)
    assert p == pytest.approx(fc * 0.5, rel=1e-14)


# --- default_equilibration_check_impl ------------------------------------------------
#
# Notation for the traces: R = N - start1 values remain; the first partition is
# [start1, start2), the second [start2, N); the second starts at N/2 for even
# N and N/2 + 1 for odd N, so it holds floor(R/2) values and each move drops
# obs[start1] (and, when R is even, moves obs[start2] into the first).

def test_equilibration_odd_ramp_plateau():
    """N = 7, obs = 5 5 5 0 2 1 0, prec 0.5.
    start1 0: means 15/4 = 3.75 vs 3/3 = 1 -> move (R odd: drop obs[0]).
    start1 1: 10/3 vs 1 -> move (R even: drop obs[1], obs[4] crosses).
    start1 2: (5+0+2)/3 = 7/3 vs (1+0)/2 -> move.
    start1 3: (0+2)/2 = 1 vs 0.5: |diff| = 0.5, not > 0.5 -> stop.
    Crossing: total mean (2+1)/4 = 0.75; obs[3] = 0 is below, obs[4] = 2 is not -> 4."""
    assert O.equilibration([5, 5, 5, 0, 2, 1, 0], 0.5) == (True, 4)


def test_equilibration_odd_never_settles():
    """N = 7, obs = 6 5 1 0 0 2 0, prec 0.5.
    start1 0: 12/4 vs 2/3; 1: 6/3 vs 2/3; 2: 1/3 vs 2/2; 3: 0/2 vs 2/2;
    4: (0+2)/2 vs 0/1 -- every |diff| > 0.5 -> move; start1 5 = N - 2 stops the scan.
    Crossing: total mean (2+0)/2 = 1; obs[5] = 2 is above, advance to 6 = N - 1,
    where the advance stops: not equilibrated, 6 samples."""
    assert O.equilibration([6, 5, 1, 0, 0, 2, 0], 0.5) == (False, 6)


def test_equilibration_even_ramp_plateau():
    """N = 8, obs = 6 6 5 1 0 1 0 0, prec 0.5.
    start1 0: 18/4 vs 1/4; 1: (6+5+1+0)/4 = 3 vs 1/3; 2: (5+1+0)/3 = 2 vs 1/3;
    3: (1+0+1)/3 vs 0/2 = 0 -> 2/3 > 0.5, move; 4: (0+1)/2 = 0.5 vs 0 -> stop.
    Crossing: total mean 1/4; obs[4] = 0 is below, obs[5] = 1 is not -> 5."""
    assert O.equilibration([6, 6, 5, 1, 0, 1, 0, 0], 0.5) == (True, 5)


def test_equilibration_difference_equal_to_prec_stops():
    """N = 3, obs = 0 3 0, prec 1.5: means (0+3)/2 = 1.5 vs 0 differ by exactly
    1.5, which does not exceed prec -> the scan stops at 0.  Crossing: total mean
    1; obs[0] = 0 is below, obs[1] = 3 is not -> 1 (equilibrated).  With ">="
    the scan would move on and end at 2 = N - 1 (not equilibrated)."""
    assert O.equilibration([0, 3, 0], 1.5) == (True, 1)


def test_equilibration_even_crossing_uses_the_remaining_mean():
    """N = 4, obs = 0 0 3 1, prec 0.5.
    start1 0: (0+0)/2 = 0 vs (3+1)/2 = 2 -> move (R even: obs[2] crosses);
    start1 1: (0+3)/2 = 1.5 vs 1 -> 0.5, stop.  Crossing: the mean of
    obs[1..3] is 4/3; obs[1] = 0 is below, obs[2] = 3 is not -> 2.
    (The whole series' mean 1 would give the same 2; the first move dropping
    only obs[0], the odd-N move, would end at 3 = N - 1.)"""
    assert O.equilibration([0, 0, 3, 1], 0.5) == (True, 2)


def test_equilibration_strict_ramp_never_equilibrates():
    """A strictly increasing ramp: the two partition means differ by about R/2
    at every start, so the scan runs to N - 2; the crossing step then reaches
    N - 1 and the series is not equilibrated."""
    assert O.equilibration(list(range(10)), 0.1) == (False, 9)
    assert O.equilibration(list(range(9)), 0.1) == (False, 8)


def test_equilibration_crossing_mean_of_the_window():
    """N = 3, obs = 0 1 2, prec 0.5: start1 0: (0+1)/2 vs 2 -> move (R odd: drop
    obs[0]); start1 1 = N - 2 stops.  The mean of obs[1..2] is 1.5: obs[1] = 1 is
    below, advance to 2 = N - 1 -> not equilibrated.  The whole series' mean 1
    would leave start1 at 1 (equilibrated)."""
    assert O.equilibration([0, 1, 2], 0.5) == (False, 2)


def test_equilibration_two_values():
    """N = 2, obs = 1 4: the scan cannot move (start1 0 = N - 2); the mean is 2.5,
    obs[0] = 1 is below, advance to 1 = N - 1 -> not equilibrated."""
    assert O.equilibration([1, 4], 0.5) == (False, 1)


def test_equilibration_all_same_eps_is_relative():
    """eps = |obs[0]| * 1e-8, or 1e-8 when obs[0] == 0."""
    assert O.equilibration([1e8, 1e8 + 0.5, 1e8 - 0.75], 1e-12) == (True, 0)
    assert O.equilibration([0.0, 5e-9, -1e-8], 1e-12) == (True, 0)
    # 1e-7 above 2.0 exceeds eps = 2e-8: the scan runs (and cannot settle below 1e-12)
    assert O.equilibration([2.0, 2.0, 2.0 + 1e-7], 1e-12) != (True, 0)
    assert O.equilibration([1e8, 1e8 + 2.0, 1e8], 1e-12) != (True, 0)


# --- CompletionCheck with requested precisions (test_CompletionCheck.py:40-124) ------

@pytest.mark.parametrize("seed", [1, 2, 3, 42])
def test_completion_check_replay(seed):
    r = O.completion_check_replay(seed)
    assert r["n_samples"] >= 100
    assert r["is_complete"] == 1
    assert r["all_equilibrated"] == 1
    assert r["n_equilibration_results"] == 2
    assert r["all_converged"] == 1
    assert r["n_convergence_results"] == 2
    assert r["prec_e"] < 0.001
    assert r["prec_v"] < 0.01
    # the samples are uniform on [mean - amp/2, mean + amp/2)
    assert abs(r["mean_e"] - 1.0) < 0.05 and abs(r["mean_v"] - 20.0) < 0.5
    # every check before completion happened at a sample point 10 steps apart
    assert r["n_steps"] == 10 * r["n_samples"]
