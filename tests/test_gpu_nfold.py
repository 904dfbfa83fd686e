"""NFOLD mode on the GPU vs the oracle (refcpu::nfold_run).

The event sequence is bit-identical: per sample the integer observables, at
the end the occupation, the class-partitioned site array and the engine words.
Times and by-pass weights agree to the rounding of log1p (device vs host libm):
relative 1e-12.  Time-grid records (grid counts and times) are exact.
"""
import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O

pytestmark = pytest.mark.gpu

J = 0.1


def _run_gpu(chains, occs, passes, chunks=None, sample_period=1, time_period=0.0, reserve=0,
             options=0):
    r = cm.Run(chains, mode=cm.MODE_NFOLD, J1=J, sample_period=sample_period,
               sample_time_period=time_period, options=options)
    for i, o in enumerate(occs):
        r.set_occupation(i, o)
    if reserve:
        r.reserve_samples(reserve)
    for k in (chunks or [passes]):
        r.run(k)
    return r


def _compare(r, c, o, exact_time=False):
    rec = r.samples(c)
    assert len(rec) == len(o["S"])
    np.testing.assert_array_equal(rec["S1"], o["S"])
    np.testing.assert_array_equal(rec["M"], o["M"])
    if exact_time:
        np.testing.assert_array_equal(rec["time"], o["time"])
        np.testing.assert_array_equal(rec["weight"], o["weight"])
    else:
        np.testing.assert_allclose(rec["time"], o["time"], rtol=1e-12)
        np.testing.assert_allclose(rec["weight"], o["weight"], rtol=1e-12)
    np.testing.assert_array_equal(r.get_occupation(c), o["occ"])
    np.testing.assert_array_equal(r.occ_location(c), o["arr"])
    np.testing.assert_array_equal(r.rng_dump(c), o["engine"])


def _compare_pending(r, c, o):
    t, site, dt = r.nfold_state(c)
    assert site == o["pending_site"]
    assert t == pytest.approx(o["final_time"], rel=1e-12)
    if site >= 0:
        assert dt == pytest.approx(o["pending_dt"], rel=1e-12)


@pytest.mark.parametrize("rows,cols,T,mu,seed,passes,sp", [
    (8, 8, 2000.0, 0.0, 5, 40, 1),
    (7, 9, 1500.0, 0.1, 6, 40, 3),
    (16, 16, 900.0, -0.2, 7, 20, 1),
    (2, 9, 2500.0, 0.05, 8, 60, 1),
    (9, 2, 2100.0, -0.05, 12, 60, 1),
    (2, 2, 3000.0, 0.0, 13, 200, 1),
    (1, 12, 1800.0, 0.0, 9, 60, 2),
    (12, 1, 1400.0, 0.1, 14, 60, 1),
    (64, 64, 2200.0, 0.0, 10, 6, 1),
    (128, 128, 1200.0, 0.3, 11, 2, 1),
])
def test_nfold_by_pass_matches_oracle(rows, cols, T, mu, seed, passes, sp):
    print("case:", rows, cols, T, mu, seed, passes, sp, flush=True)
    n = rows * cols
    occ = np.where(np.random.default_rng(seed).random(n) < 0.6, 1, -1).astype(np.int8)
    o = O.nfold_run(rows, cols, J, T, mu, passes, occ, O.mt64_state(seed, 0), sample_period=sp)
    r = _run_gpu([dict(shape=(rows, cols), temperature=T, exchange_potential=mu, seed=seed)],
                 [occ], passes, chunks=[1, passes // 2, passes - 1 - passes // 2],
                 sample_period=sp)
    _compare(r, 0, o)
    _compare_pending(r, 0, o)
    n_pass, n_ev, n_rej = r.counters(0)
    assert (n_pass, n_ev, n_rej) == (passes, passes * n, 0)
    r.close()


@pytest.mark.parametrize("rows,cols,seed,passes,sp", [(8, 8, 21, 30, 1), (5, 7, 22, 31, 3)])
def test_nfold_sample_start_and_total_rate(rows, cols, seed, passes, sp):
    """MCG_OPT_NFOLD_SAMPLE_START adds the pass-0 record (the initial state
    weighted by the first event's increment); by-pass records keep the total
    rate R in e_run."""
    print("case:", rows, cols, seed, passes, sp, flush=True)
    T, mu = 1700.0, 0.05
    occ = np.where(np.random.default_rng(seed).random(rows * cols) < 0.5, 1, -1).astype(np.int8)
    o = O.nfold_run(rows, cols, J, T, mu, passes, occ, O.mt64_state(seed, 0), sample_period=sp,
                    sample_start=True)
    r = _run_gpu([dict(shape=(rows, cols), temperature=T, exchange_potential=mu, seed=seed)],
                 [occ], passes, chunks=[0, 0, 1, passes - 1], sample_period=sp,
                 options=cm.OPT_NFOLD_SAMPLE_START)
    rec = r.samples(0)
    assert len(rec) == passes // sp + 1
    assert rec["pass"][0] == 0 and rec["time"][0] == 0.0
    _compare(r, 0, o)
    _compare_pending(r, 0, o)
    np.testing.assert_allclose(rec["e_run"], o["rate"], rtol=1e-13)
    r.close()


def test_nfold_zero_pass_run_takes_the_start_sample():
    """mcg_run(h, 0) with MCG_OPT_NFOLD_SAMPLE_START records pass 0 and leaves
    the first event pending; the run that follows applies it first."""
    rows, cols, seed, T, mu = 6, 10, 23, 1900.0, -0.05
    occ = np.where(np.random.default_rng(seed).random(rows * cols) < 0.5, 1, -1).astype(np.int8)
    o0 = O.nfold_run(rows, cols, J, T, mu, 0, occ, O.mt64_state(seed, 0), sample_start=True)
    r = _run_gpu([dict(shape=(rows, cols), temperature=T, exchange_potential=mu, seed=seed)],
                 [occ], 0, chunks=[0], options=cm.OPT_NFOLD_SAMPLE_START)
    assert r.counters(0) == (0, 0, 0)
    _compare(r, 0, o0)
    _compare_pending(r, 0, o0)
    assert o0["pending_site"] >= 0
    r.run(0)  # no further sample: the first event is already pending
    assert r.n_samples(0) == 1
    r.run(5)
    o = O.nfold_run(rows, cols, J, T, mu, 5, occ, O.mt64_state(seed, 0), sample_start=True)
    _compare(r, 0, o)
    _compare_pending(r, 0, o)
    r.close()


def test_nfold_record_on_demand():
    """With no periodic records, mcg_run(h, 0) after pass 3 records that pass
    and selects its first event (a second call adds nothing); the trajectory is
    the oracle's with a sample at pass 3 only."""
    rows, cols, seed, T, mu = 7, 8, 24, 2100.0, 0.1
    occ = np.where(np.random.default_rng(seed).random(rows * cols) < 0.4, 1, -1).astype(np.int8)
    o = O.nfold_run(rows, cols, J, T, mu, 5, occ, O.mt64_state(seed, 0), sample_period=3)
    assert len(o["S"]) == 1
    r = _run_gpu([dict(shape=(rows, cols), temperature=T, exchange_potential=mu, seed=seed)],
                 [occ], 0, chunks=[3, 0, 0, 2], sample_period=1 << 40)
    rec = r.samples(0)
    assert len(rec) == 1 and rec["pass"][0] == 3
    _compare(r, 0, o)
    _compare_pending(r, 0, o)
    assert r.counters(0) == (5, 5 * rows * cols, 0)
    r.close()


def test_nfold_time_grid_matches_oracle():
    rows, cols, T, mu, seed, passes = 12, 10, 1600.0, 0.05, 21, 30
    occ = np.ones(rows * cols, dtype=np.int8)
    period = O.nfold_run(rows, cols, J, T, mu, passes, occ, O.mt64_state(seed, 0),
                         time_period=1e30, capacity=1)["final_time"] / 400
    o = O.nfold_run(rows, cols, J, T, mu, passes, occ, O.mt64_state(seed, 0),
                    time_period=period, capacity=10000)
    r = _run_gpu([dict(shape=(rows, cols), temperature=T, exchange_potential=mu, seed=seed)],
                 [occ], passes, chunks=[7, passes - 7], time_period=period, reserve=10000)
    _compare(r, 0, o, exact_time=True)
    r.close()


def test_nfold_time_grid_batch():
    """Several chains on the time grid, a grid fine enough that most residence
    intervals hold several grid points and coarse enough that many hold none."""
    shapes = [(10, 10), (7, 3), (16, 8)]
    chains, occs, outs = [], [], []
    rng = np.random.default_rng(44)
    for i, (a, b) in enumerate(shapes):
        T, mu, seed = 900.0 + 700.0 * i, 0.1 - 0.1 * i, 800 + i
        occ = np.where(rng.random(a * b) < 0.5, 1, -1).astype(np.int8)
        chains.append(dict(shape=(a, b), temperature=T, exchange_potential=mu, seed=seed))
        occs.append(occ)
    period = 0.02
    for i, (a, b) in enumerate(shapes):
        outs.append(O.nfold_run(a, b, J, chains[i]["temperature"],
                                chains[i]["exchange_potential"], 9, occs[i],
                                O.mt64_state(chains[i]["seed"], 0), time_period=period,
                                capacity=20000))
    r = _run_gpu(chains, occs, 9, chunks=[4, 5], time_period=period, reserve=20000)
    for i in range(len(shapes)):
        assert len(outs[i]["S"]) > 10
        _compare(r, i, outs[i], exact_time=True)
    r.close()


def test_nfold_frozen_time_grid():
    """All +1 at 1 K with mu = 0.4 eV: every flip raises the potential by 1.2 eV,
    its Metropolis probability underflows to 0, the total rate is 0 and the
    next event never comes; the time grid stops with an error."""
    r = cm.Run([dict(shape=(8, 8), temperature=1.0, exchange_potential=0.4, seed=1)],
               mode=cm.MODE_NFOLD, J1=J, sample_time_period=1e-3)
    with pytest.raises(cm.McgError, match="total rate is 0"):
        r.run(1)
    r.close()


def test_nfold_batch():
    rng = np.random.default_rng(33)
    shapes = [(8, 8), (9, 7), (20, 12), (5, 5), (32, 16), (6, 30)]
    chains, occs = [], []
    for i, (a, b) in enumerate(shapes):
        chains.append(dict(shape=(a, b), temperature=700.0 + 500.0 * i,
                           exchange_potential=-0.3 + 0.12 * i, seed=500 + i))
        occs.append(np.where(rng.random(a * b) < 0.5, 1, -1).astype(np.int8))
    r = _run_gpu(chains, occs, 12)
    for i, ((a, b), occ) in enumerate(zip(shapes, occs)):
        o = O.nfold_run(a, b, J, chains[i]["temperature"], chains[i]["exchange_potential"], 12,
                        occ, O.mt64_state(500 + i, 0))
        _compare(r, i, o)
    r.close()


def test_nfold_errors():
    r = _run_gpu([dict(shape=(6, 6), temperature=1500.0, seed=1)],
                 [np.ones(36, dtype=np.int8)], 2)
    with pytest.raises(cm.McgError, match="invalid method"):
        r.stats_weighted(0, cm.OBS_POTENTIAL_ENERGY, method=3)
    with pytest.raises(cm.McgError, match="n_resamples"):
        r.stats_weighted(0, cm.OBS_POTENTIAL_ENERGY, n_resamples=0)
    r.close()
    e = cm.Run([dict(shape=(6, 6), temperature=1500.0, seed=1)], mode=cm.MODE_EXACT, J1=J)
    e.run(2)
    with pytest.raises(cm.McgError, match="weighted records"):
        e.stats_weighted(0, cm.OBS_POTENTIAL_ENERGY)
    e.close()
    r = cm.Run([dict(shape=(6, 6), temperature=5000.0, seed=1)], mode=cm.MODE_NFOLD, J1=J,
               sample_time_period=1e-6)
    r.reserve_samples(10)
    with pytest.raises(cm.McgError, match="exceeded"):
        r.run(2)
    r.close()


def test_nfold_at_the_site_limit():
    """181 x 181 = 32761 sites (positions need 15 bits; bit 15 is the spin):
    bit-exact for one pass; 182 x 181 is refused."""
    rows, cols, T, mu, seed = 181, 181, 1700.0, 0.05, 61
    occ = np.where(np.random.default_rng(seed).random(rows * cols) < 0.5, 1, -1).astype(np.int8)
    o = O.nfold_run(rows, cols, J, T, mu, 1, occ, O.mt64_state(seed, 0))
    r = _run_gpu([dict(shape=(rows, cols), temperature=T, exchange_potential=mu, seed=seed)],
                 [occ], 1)
    _compare(r, 0, o)
    r.close()
    with pytest.raises(cm.McgError, match="NFOLD at most 32768"):
        cm.Run([dict(shape=(182, 181), temperature=T, seed=1)], mode=cm.MODE_NFOLD, J1=J)
