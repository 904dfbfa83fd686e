"""Weighted statistics of NFOLD records on the GPU vs the oracle
(refcpu::weighted_statistics, BasicStatistics.cc:144-188) and the weighted
equilibration check (EquilibrationCheck.cc:119-162).

Observables are formed from the records' integer S1 and M exactly as the
device forms them; the device sums in another order than the oracle's
sequential loops, so mean and variance agree to 1e-12 (relative) and the
autocorrelation factor and precision to 1e-9.  On the time grid every weight
is an integer, so the total weight and the resampling points are exact.  The
weighted equilibration scan is sequential on both sides and must agree
exactly.
"""
import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O

pytestmark = pytest.mark.gpu

J = 0.1


def _obs(rec, n, observable, mu):
    S1 = rec["S1"].astype(np.float64)
    Nx = (n + rec["M"].astype(np.float64)) / 2.0
    ef = S1 * (-J)
    if observable == cm.OBS_FORMATION_ENERGY:
        return ef / n
    if observable == cm.OBS_PARAM_COMPOSITION:
        return Nx / n
    return (ef - mu * Nx) / n


def _run(shape, T, mu, seed, passes, time_period=0.0):
    n = shape[0] * shape[1]
    r = cm.Run([dict(shape=shape, temperature=T, exchange_potential=mu, seed=seed)],
               mode=cm.MODE_NFOLD, J1=J, sample_time_period=time_period)
    occ = np.where(np.random.default_rng(seed).random(n) < 0.5, 1, -1).astype(np.int8)
    r.set_occupation(0, occ)
    r.reserve_samples(60000)
    r.run(passes)
    return r, n


def _check(r, n, mu, first, exact_weights):
    rec = r.samples(0)[first:]
    w = rec["weight"]
    for observable in (cm.OBS_FORMATION_ENERGY, cm.OBS_PARAM_COMPOSITION, cm.OBS_POTENTIAL_ENERGY):
        x = _obs(rec, n, observable, mu)
        for method, nres in ((1, 10000), (2, 4000)):
            print("observable", observable, "method", method, flush=True)
            g = r.stats_weighted(0, observable, first=first, method=method, n_resamples=nres)
            o = O.weighted_statistics(x, w, 0.95, method, nres)
            assert g["n"] == len(x)
            assert g["mean"] == pytest.approx(o["mean"], rel=1e-12, abs=1e-15)
            assert g["variance"] == pytest.approx(o["variance"], rel=1e-10, abs=1e-18)
            tol = 1e-12 if exact_weights else 1e-9
            assert g["autocorr_factor"] == pytest.approx(o["autocorr_factor"], rel=tol)
            assert g["calculated_precision"] == pytest.approx(o["calculated_precision"], rel=1e-9)
        # mcg_stats_compute on NFOLD records is method 1 with 10000 resamples
        d = r.stats(0, observable, first=first)
        m1 = r.stats_weighted(0, observable, first=first)
        assert d == m1


def test_weighted_stats_by_pass():
    r, n = _run((9, 7), 1400.0, 0.05, 31, 800)
    _check(r, n, 0.05, 100, exact_weights=False)
    r.close()


def test_weighted_stats_time_grid():
    shape, T, mu, seed = (10, 10), 1800.0, -0.1, 32
    n = 100
    period = O.nfold_run(10, 10, J, T, mu, 300, np.where(
        np.random.default_rng(seed).random(n) < 0.5, 1, -1).astype(np.int8),
        O.mt64_state(seed, 0), time_period=1e30, capacity=1)["final_time"] / 20000
    r, n = _run(shape, T, mu, seed, 300, time_period=period)
    assert np.all(r.samples(0)["weight"] == np.round(r.samples(0)["weight"]))
    _check(r, n, mu, 0, exact_weights=True)
    r.close()


@pytest.mark.parametrize("prec", [1e-3, 1e-4])
def test_weighted_equilibration_exact(prec):
    r, n = _run((9, 7), 900.0, 0.1, 33, 600)
    rec = r.samples(0)
    w = rec["weight"]
    W = 0.0
    for v in w:  # in order, as the reference's loop
        W += float(v)
    wf = len(w) / W
    for observable in (cm.OBS_FORMATION_ENERGY, cm.OBS_PARAM_COMPOSITION, cm.OBS_POTENTIAL_ENERGY):
        x = _obs(rec, n, observable, 0.1)
        xw = x * (wf * w)
        assert r.equilibration(0, observable, prec) == O.equilibration(xw, prec)
    r.close()
