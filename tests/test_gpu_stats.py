"""Device statistics over device-resident sample series vs the oracle.

mcg_stats_compute (BasicStatisticsCalculator, BasicStatistics.cc:114-131) and
mcg_equilibration (_default_equilibration_check, EquilibrationCheck.cc:50-117)
reduce the series on the GPU in a different summation order than the
reference's sequential loops, so the bar is relative 1e-12 on mean and
precision and equal decisions (lag, equilibration index) except at exact ties.
"""
import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O

pytestmark = pytest.mark.gpu


def check_stats(r, c, obs_id, x):
    m, p = O.basic_statistics(x)
    s = r.stats(c, obs_id)
    assert s["n"] == len(x)
    assert s["mean"] == pytest.approx(m, rel=1e-12, abs=1e-15)
    assert s["calculated_precision"] == pytest.approx(p, rel=1e-9, abs=1e-15)
    assert s["autocorr_factor"] == pytest.approx(O.autocorrelation_factor(x), rel=1e-12)
    for prec in (1e-3, 1e-2):
        is_eq, n_eq = r.equilibration(c, obs_id, prec)
        assert (is_eq, n_eq) == O.equilibration(x, prec)


@pytest.mark.parametrize("T,mu", [(2000.0, 0.0), (3500.0, 0.1), (900.0, -0.2)])
def test_exact_series_stats(T, mu):
    print("case: exact stats", T, mu, flush=True)
    r = cm.Run([dict(shape=(25, 25), temperature=T, exchange_potential=mu, seed=17)],
               mode=cm.MODE_EXACT, J1=0.1)
    r.run(3000)
    obs = r.observables(0)
    for obs_id, name in ((cm.OBS_FORMATION_ENERGY, "formation_energy"),
                         (cm.OBS_PARAM_COMPOSITION, "param_composition"),
                         (cm.OBS_POTENTIAL_ENERGY, "potential_energy")):
        check_stats(r, 0, obs_id, obs[name])


def test_checker_series_stats_batched():
    """Statistics of every chain of a batch, including J2 terms."""
    chains = [dict(shape=(32, 32), temperature=1500.0 + 400 * i, exchange_potential=0.05 * i,
                   seed=70 + i) for i in range(6)]
    print("case: checker stats batched", flush=True)
    r = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, J1=0.1, J2=-0.02)
    r.run(4000)
    for c in range(len(chains)):
        obs = r.observables(c)
        check_stats(r, c, cm.OBS_POTENTIAL_ENERGY, obs["potential_energy"])
        check_stats(r, c, cm.OBS_PARAM_COMPOSITION, obs["param_composition"])


def test_stats_edge_cases():
    """A constant series (all samples equal) and a two-sample series."""
    r = cm.Run([dict(shape=(4, 4), temperature=50.0, exchange_potential=-1.0, seed=3)],
               mode=cm.MODE_EXACT, J1=0.1)
    r.run(200)  # frozen in the all -1 state after the first passes
    x = r.observables(0)["param_composition"]
    s = r.stats(0, cm.OBS_PARAM_COMPOSITION, first=100)
    m, p = O.basic_statistics(x[100:])
    assert (s["mean"], s["calculated_precision"]) == (m, p)
    assert r.equilibration(0, cm.OBS_PARAM_COMPOSITION, 1e-3) == O.equilibration(x, 1e-3)
    s2 = r.stats(0, cm.OBS_POTENTIAL_ENERGY, first=198)
    assert s2["n"] == 2
    with pytest.raises(cm.McgError):
        r.stats(0, cm.OBS_POTENTIAL_ENERGY, first=200)  # no observations


@pytest.mark.parametrize("mode", [cm.MODE_EXACT, cm.MODE_CHECKERBOARD])
def test_check_batch_equals_single_chain_checks(mode):
    """mcg_check_batch (one launch for many chains) gives every chain the
    equilibration index and tail statistics of mcg_equilibration and
    mcg_stats_compute bit for bit, in any chain order."""
    print("case: check_batch mode", mode, flush=True)
    conds = [dict(shape=(16, 16), temperature=1200.0 + 400.0 * k,
                  exchange_potential=-0.2 + 0.1 * k, seed=70 + k) for k in range(6)]
    r = cm.Run(conds, mode=mode, J1=0.1)
    r.run(1500)
    chains = [4, 0, 5, 2, 1, 3]
    for obs in (cm.OBS_PARAM_COMPOSITION, cm.OBS_POTENTIAL_ENERGY):
        for prec in (1e-3, 1e-2):
            got = r.check_batch(chains, obs, prec)
            for c, g in zip(chains, got):
                is_eq, n_eq = r.equilibration(c, obs, prec)
                assert (g["is_equilibrated"], g["n_equilibration"]) == (is_eq, n_eq)
                if is_eq and n_eq < r.n_samples(c):
                    s = r.stats(c, obs, first=n_eq)
                    assert g["n"] == s["n"]
                    for k in ("mean", "variance", "autocorr_factor", "calculated_precision"):
                        assert np.float64(g[k]).view(np.uint64) == \
                            np.float64(s[k]).view(np.uint64), k
                else:
                    assert g["n"] == 0
    r.close()
