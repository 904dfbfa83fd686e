"""Fluctuation observables of a sampled SGC run (heat capacity and the
parametric-composition susceptibility) and the weighted mean of a weighted
(N-fold way) series, with block error bars.

The reference names these analyses ("heat_capacity", "param_susc",
run_management/io/json/SamplingFixtureParams_json_io.hh:54-59) but implements
them in libcasm-clexmonte, which is not part of libcasm-monte; the
conventions here are the textbook fluctuation formulas, per unit cell:
  C     = N * var(e) / (kB T^2),   e = potential energy per unit cell,
  chi_x = N * var(x) / (kB T),     x = parametric composition,
where var is the population variance of the series (N = unit cells).
Errors, two ways:
  * block errors (heat_capacity, susceptibility): the series is cut into
    `blocks` consecutive blocks, the quantity is evaluated per block, and
    sigma = std(block values) / sqrt(blocks);
  * the BasicStatistics convention (heat_capacity_stats, susceptibility_stats):
    var(e) is the mean of the squared deviations y_i = (e_i - mean(e))^2, so
    its precision is BasicStatistics' calculated_precision of the series y
    (BasicStatistics.cc:114-131: the confidence half-width with the
    autocorrelation factor of BasicStatistics.cc:24-48), scaled like var.
"""
import math

import numpy as np

from . import KB


def approx_erf_inv(x):
    """Winitzki's approximation of erf^-1 with a = 0.147 (misc/math.hh:62-72)."""
    a = 0.147
    sgn = -1.0 if x < 0.0 else 1.0
    b = math.log((1.0 - x) * (1.0 + x))
    c = 2.0 / (math.pi * a) + b * 0.5
    d = b / a
    return sgn * math.sqrt(math.sqrt(c * c - d) - c)


def autocorrelation_factor(x):
    """(1 + rho) / (1 - rho), rho = 2^(-1/k) for the first lag k with
    |cov_k / var0| <= 0.5, cov_k = sum_{j < N-k} (x_j - m)(x_{j+k} - m) / (N - k);
    1 when |var0 / mean| < 1e-8 or var0 == 0; DBL_MAX when no lag qualifies
    (BasicStatistics.cc:24-48)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    m = x.mean()
    d = x - m
    v0 = float(np.dot(d, d)) / n
    if v0 == 0.0 or abs(v0 / m) < 1e-8:
        return 1.0
    for k in range(1, n):
        cov = float(np.dot(d[:n - k], d[k:])) / (n - k)
        if abs(cov / v0) <= 0.5:
            rho = 2.0 ** (-1.0 / k)
            return (1.0 + rho) / (1.0 - rho)
    return float(np.finfo(np.float64).max)


def basic_statistics(x, confidence=0.95):
    """(mean, calculated_precision) of a series: sqrt(2) erf^-1(confidence) *
    sqrt(f * var0 / N) with f the autocorrelation factor
    (BasicStatistics.cc:114-131).  Sums are numpy's (pairwise), so results
    agree with a sequential sum to rounding."""
    x = np.asarray(x, dtype=np.float64)
    if len(x) == 0:
        raise ValueError("observations.size()==0")
    m = float(x.mean())
    v0 = float(np.mean((x - m) ** 2))
    f = autocorrelation_factor(x)
    fc = math.sqrt(2.0) * approx_erf_inv(confidence)
    return m, fc * math.sqrt(f * v0 / len(x))


def _variance_stats(x, confidence):
    x = np.asarray(x, dtype=np.float64)
    y = (x - x.mean()) ** 2
    return basic_statistics(y, confidence)


def heat_capacity_stats(e_per_unitcell, n_unitcells, temperature, kb=KB, confidence=0.95):
    """C = N var(e) / (kB T^2) and its calculated precision (the BasicStatistics
    half-width at `confidence` of var(e), scaled the same way)."""
    v, p = _variance_stats(e_per_unitcell, confidence)
    s = n_unitcells / (kb * temperature ** 2)
    return v * s, p * s


def susceptibility_stats(x, n_unitcells, temperature, kb=KB, confidence=0.95):
    """chi_x = N var(x) / (kB T) and its calculated precision (as heat_capacity_stats)."""
    v, p = _variance_stats(x, confidence)
    s = n_unitcells / (kb * temperature)
    return v * s, p * s


def _blocked(x, fn, blocks):
    x = np.asarray(x, dtype=np.float64)
    n = len(x) // blocks
    vals = np.array([fn(x[i * n:(i + 1) * n]) for i in range(blocks)])
    return float(fn(x)), float(vals.std(ddof=1) / np.sqrt(blocks))


def heat_capacity(e_per_unitcell, n_unitcells, temperature, kb=KB, blocks=20):
    f = lambda s: n_unitcells * np.var(s) / (kb * temperature ** 2)
    return _blocked(e_per_unitcell, f, blocks)


def susceptibility(x, n_unitcells, temperature, kb=KB, blocks=20):
    f = lambda s: n_unitcells * np.var(s) / (kb * temperature)
    return _blocked(x, f, blocks)


def weighted_mean(x, w, blocks=20):
    """Weighted mean sum(x w) / sum(w) of a weighted series (NFOLD samples: w are
    residence times or time-grid counts) with a block error: the series is cut
    into `blocks` consecutive blocks of equal length and
    sigma = std(block weighted means) / sqrt(blocks)."""
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    mean = float((x * w).sum() / w.sum())
    n = len(x) // blocks
    if n == 0:
        return mean, float("nan")
    xb = (x[:n * blocks] * w[:n * blocks]).reshape(blocks, n).sum(1)
    wb = w[:n * blocks].reshape(blocks, n).sum(1)
    return mean, float((xb / wb).std(ddof=1) / np.sqrt(blocks))
