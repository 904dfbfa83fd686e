"""Pins of the oracle's N-fold way chain (refcpu::nfold_run).

The reference's nfold loop (nfold.hh:78-149) takes its event selector from
outside the repository, so the selector is defined here (class-based
rates, one partitioned site array); the pins are:
  * the partition: the site array holds every site once, grouped by class in
    class order, after runs;
  * stream use: exactly three engine words per selected event;
  * the time-weighted averages (weights = residence times) of E and N x equal
    the SGC ensemble averages from exact enumeration of a 4x4 lattice, and
    mean residence time 1/N at infinite temperature (all rates 1);
  * the integer observables equal those recomputed from the final state.
"""
import numpy as np
import pytest

import oracle

J = 0.1


def _class_of(s, rows, cols, l):
    i, j = l % rows, l // rows
    nb = (s[(i + 1) % rows + rows * j] + s[(i - 1) % rows + rows * j] +
          s[i + rows * ((j + 1) % cols)] + s[i + rows * ((j - 1) % cols)])
    return (5 if s[l] > 0 else 0) + (nb + 4) // 2


def _bond_sum(s, rows, cols):
    g = s.reshape(cols, rows).T.astype(np.int64)
    return int((g * np.roll(g, -1, 0)).sum() + (g * np.roll(g, -1, 1)).sum())


@pytest.mark.parametrize("rows,cols,T,mu", [(8, 8, 2000.0, 0.0), (6, 10, 900.0, 0.2),
                                           (5, 7, 3500.0, -0.3)])
def test_partition_and_observables(rows, cols, T, mu):
    n = rows * cols
    occ = np.where(np.random.default_rng(rows).random(n) < 0.5, 1, -1).astype(np.int8)
    r = oracle.nfold_run(rows, cols, J, T, mu, 30, occ, oracle.mt64_state(3, 0))
    s = r["occ"]
    assert sorted(r["arr"].tolist()) == list(range(n))
    classes = [_class_of(s, rows, cols, l) for l in r["arr"]]
    assert classes == sorted(classes)
    assert r["S"][-1] == _bond_sum(s, rows, cols) and r["M"][-1] == int(s.sum())
    assert np.all(np.diff(r["time"]) > 0) and np.all(r["weight"] > 0)


def test_three_words_per_event():
    rows = cols = 6
    seed = 11
    r = oracle.nfold_run(rows, cols, J, 2500.0, 0.0, 3, np.ones(36, dtype=np.int8),
                         oracle.mt64_state(seed, 0))
    events = 3 * 36 + 1  # three passes and the selected event the last sample weighs
    assert np.array_equal(oracle.mt64_state(seed, 3 * events), r["engine"])


def test_infinite_temperature_residence_time():
    """All rates 1: R = N and dt = -log(1-u)/N, mean 1/N."""
    n = 64
    r = oracle.nfold_run(8, 8, J, 1e12, 0.0, 4000, np.ones(n, dtype=np.int8),
                         oracle.mt64_state(7, 0))
    w = r["weight"]
    assert abs(w.mean() * n - 1.0) < 4 * w.std() * n / np.sqrt(len(w))
    assert abs(r["final_time"] / (4000 * n) * n - 1.0) < 0.01


def _check_vs_enumeration(r, rows, cols, T, mu, drop):
    n = rows * cols
    ex = oracle.lattice_enumerate(2, (rows, cols), J, 0.0, T, mu)
    S, M, w = r["S"][drop:], r["M"][drop:], r["weight"][drop:]
    Ef = -J * S
    Nx = (n + M) / 2.0
    E = Ef - mu * Nx  # SGC potential per supercell
    nb = 100
    m = len(w) // nb * nb
    for x, exact in ((E, ex["E"]), (Nx, ex["Nx"])):
        xb = (x[:m] * w[:m]).reshape(nb, -1).sum(1) / w[:m].reshape(nb, -1).sum(1)
        mean = (x * w).sum() / w.sum()
        err = xb.std(ddof=1) / np.sqrt(nb)
        assert abs(mean - exact) < 4 * err + 1e-9, (mean, exact, err)


@pytest.mark.parametrize("T,mu", [(2000.0, 0.0), (1500.0, 0.15)])
def test_time_grid_sampling_vs_enumeration(T, mu):
    """Samples on a time grid (one record per residence interval, weighted by
    its number of grid points) give the ensemble averages (4x4)."""
    args = (4, 4, J, T, mu, 60000, np.ones(16, dtype=np.int8), oracle.mt64_state(2024, 0))
    period = oracle.nfold_run(*args, time_period=1e30, capacity=1)["final_time"] / 500000
    r = oracle.nfold_run(*args, time_period=period, capacity=600_000)
    w, t = r["weight"], r["time"]
    assert r["n_total"] == len(r["S"]) and np.all(w >= 1.0) and np.all(w == np.round(w))
    # records are runs of consecutive grid points: k_next = k + weight
    k = np.round(t / period)
    assert np.allclose(t, k * period) and np.array_equal(k[1:], k[:-1] + w[:-1])
    _check_vs_enumeration(r, 4, 4, T, mu, 1000)


def test_pass_sampling_odd_lattice_vs_enumeration():
    """By-pass samples weighted by residence times give the ensemble averages
    when N is odd (3x5): consecutive pass samples alternate flip parity."""
    r = oracle.nfold_run(3, 5, J, 1800.0, 0.05, 60000, np.ones(15, dtype=np.int8),
                         oracle.mt64_state(99, 0))
    _check_vs_enumeration(r, 3, 5, 1800.0, 0.05, 500)


def test_pass_sampling_even_lattice_sees_one_parity():
    """With N even every pass sample has an even number of flips since the
    start: the number of -1 sites keeps the initial parity (this is why the
    time grid is the sampling to use for even N)."""
    r = oracle.nfold_run(4, 4, J, 2000.0, 0.0, 2000, np.ones(16, dtype=np.int8),
                         oracle.mt64_state(5, 0))
    n_down = (16 - r["M"]) // 2
    assert np.all(n_down % 2 == 0)


def test_weighted_mean_helper():
    """analysis.weighted_mean on a by-pass series equals the direct weighted sum
    and agrees with enumeration (odd N, 3x5)."""
    from casmcode_monte_amd import analysis
    r = oracle.nfold_run(3, 5, J, 1800.0, 0.05, 20000, np.ones(15, dtype=np.int8),
                         oracle.mt64_state(4, 0))
    Nx = (15 + r["M"][200:]) / 2.0
    w = r["weight"][200:]
    mean, err = analysis.weighted_mean(Nx, w, blocks=50)
    assert mean == pytest.approx((Nx * w).sum() / w.sum(), rel=1e-15)
    ex = oracle.lattice_enumerate(2, (3, 5), J, 0.0, 1800.0, 0.05)
    assert abs(mean - ex["Nx"]) < 4 * err


def test_sample_start_and_total_rate():
    """The count-0 sample holds the initial state at time 0, weighted by the first
    event's increment; selecting that event early changes no later draw, so the
    rest equals the run without it.  On an all +1 lattice every site is in the
    class (s = +1, four +1 neighbours), so the recorded total rate is
    N * exp(-beta * (8 J + mu))."""
    rows, cols, T, mu, seed = 6, 7, 1800.0, 0.1, 12
    occ = np.ones(rows * cols, dtype=np.int8)
    a = oracle.nfold_run(rows, cols, 0.1, T, mu, 5, occ, oracle.mt64_state(seed, 0),
                         sample_start=True)
    b = oracle.nfold_run(rows, cols, 0.1, T, mu, 5, occ, oracle.mt64_state(seed, 0))
    assert len(a["S"]) == len(b["S"]) + 1
    assert a["S"][0] == 2 * rows * cols and a["M"][0] == rows * cols and a["time"][0] == 0.0
    np.testing.assert_array_equal(a["S"][1:], b["S"])
    np.testing.assert_array_equal(a["weight"][1:], b["weight"])
    np.testing.assert_array_equal(a["occ"], b["occ"])
    np.testing.assert_array_equal(a["engine"], b["engine"])
    beta = 1.0 / (oracle.kb() * T)
    assert a["rate"][0] == pytest.approx(rows * cols * np.exp(-beta * (8 * 0.1 + mu)), rel=1e-13)
    # the first increment is -log1p(-u)/R with u the canonical double of the
    # third word the first event draws
    u = oracle.canonical(oracle.mt64_outputs(seed, 3)[2])
    assert a["weight"][0] == -np.log1p(-u) / a["rate"][0]
