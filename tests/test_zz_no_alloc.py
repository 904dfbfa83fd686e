"""mcg_stats_compute_noalloc (the statistics entry point that allocates nothing)
against mcg_stats_compute, through the C ABI.

Both run the same kernels on the same device series, so the results are
bit-identical; the no-allocation entry point computes in the library's fixed
device scratch instead of the handle's growable one.  Checked on two calls in a
row, the second on a smaller input (a later first sample), for the unweighted
branch (EXACT and CHECKERBOARD records) and the weighted one (NFOLD records,
10000 resamples), and the refusal of a series longer than the scratch.
"""
import numpy as np
import pytest

import casmcode_monte_amd as cm

pytestmark = pytest.mark.gpu

KEYS = ("mean", "calculated_precision", "variance", "autocorr_factor")


def same(a, b):
    assert a["n"] == b["n"]
    for k in KEYS:
        assert np.float64(a[k]).view(np.uint64) == np.float64(b[k]).view(np.uint64), k


def test_noalloc_equals_stats_compute_exact():
    print("case: noalloc exact 25x25 x 3000 passes", flush=True)
    r = cm.Run([dict(shape=(25, 25), temperature=2000.0, exchange_potential=0.0, seed=17),
                dict(shape=(16, 20), temperature=3200.0, exchange_potential=0.1, seed=18)],
               mode=cm.MODE_EXACT, J1=0.1)
    r.run(3000)
    for c in (0, 1):
        for first in (0, 2400):  # the second call reads a smaller input
            for obs in (cm.OBS_FORMATION_ENERGY, cm.OBS_PARAM_COMPOSITION,
                        cm.OBS_POTENTIAL_ENERGY):
                b = r.stats_noalloc(c, obs, first)
                a = r.stats(c, obs, first)
                assert a["n"] == 3000 - first
                same(a, b)
    r.close()


def test_noalloc_equals_stats_compute_checker():
    print("case: noalloc checkerboard 32x32 x 4 chains", flush=True)
    conds = [dict(shape=(32, 32), temperature=1500.0 + 400.0 * k, exchange_potential=0.0,
                  seed=40 + k) for k in range(4)]
    r = cm.Run(conds, mode=cm.MODE_CHECKERBOARD, J1=0.1, J2=-0.02)
    r.run(1500)
    for c in range(4):
        for first in (100, 1000):
            b = r.stats_noalloc(c, cm.OBS_POTENTIAL_ENERGY, first)
            a = r.stats(c, cm.OBS_POTENTIAL_ENERGY, first)
            same(a, b)
    r.close()


def test_noalloc_equals_stats_compute_weighted():
    print("case: noalloc nfold 12x12 weighted", flush=True)
    r = cm.Run([dict(shape=(12, 12), temperature=1800.0, exchange_potential=0.05, seed=5)],
               mode=cm.MODE_NFOLD, J1=0.1)
    r.run(400)
    for first in (0, 300):
        b = r.stats_noalloc(0, cm.OBS_PARAM_COMPOSITION, first)
        a = r.stats(0, cm.OBS_PARAM_COMPOSITION, first)
        same(a, b)
    r.close()


def test_noalloc_refuses_series_beyond_the_scratch():
    print("case: noalloc refusal, 2x2 x 262200 passes", flush=True)
    n = 262200  # > MCG_STATS_NOALLOC_MAX - 8 = 262136 observations
    r = cm.Run([dict(shape=(2, 2), temperature=2000.0, exchange_potential=0.0, seed=3)],
               mode=cm.MODE_EXACT, J1=0.1)
    r.reserve_samples(n)
    r.run(n)
    with pytest.raises(cm.McgError, match="fixed scratch"):
        r.stats_noalloc(0, cm.OBS_PARAM_COMPOSITION, 0)
    # the longest series it takes, then a shorter one, equal the growable path
    for first in (n - 262136, n - 1000):
        same(r.stats(0, cm.OBS_PARAM_COMPOSITION, first),
             r.stats_noalloc(0, cm.OBS_PARAM_COMPOSITION, first))
    r.close()
