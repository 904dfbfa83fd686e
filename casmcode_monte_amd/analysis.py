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
Errors: the series is cut into `blocks` consecutive blocks, the quantity is
evaluated per block, and sigma = std(block values) / sqrt(blocks).
"""
import numpy as np

from . import KB


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
