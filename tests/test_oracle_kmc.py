"""Pins of the oracle's kinetic Monte Carlo restatement (refcpu::kmc_run): the
loop of kinetic_monte_carlo.hh:136-229 with OccLocation atom tracking, for a
square lattice gas under the rejection event selector defined in
oracle/refcpu.hh (the reference takes its selector from outside).

Independent checks:
- one atom is a random walk: every attempt succeeds, the clock advances
  exactly 1/4 per event, each event draws exactly two words (atom slot and
  direction), and the squared displacement over m events averages m (m/2 per
  axis);
- the dynamics obey detailed balance for the Boltzmann distribution at fixed
  atom number, so averages over the time grid (which weights every state by
  its residence time) equal the canonical average over all configurations of
  a 3 x 4 lattice, while averages over events (by-pass samples) do not;
- with J = 0 every attempt to a vacant site is accepted without a draw;
- the OccLocation lists and the atoms stay consistent with the occupation
  after each of the first events (the invariants of the reference's
  OccLocation test, tests/unit/monte/OccLocation_test.cpp:15-65).
"""
import itertools

import numpy as np

import oracle as O

KB = 8.6173303e-05


def _bond_sum(occ, rows, cols):
    o = np.asarray(occ).reshape(cols, rows).T  # o[i, j], site l = i + rows*j
    return int((o * np.roll(o, -1, 0)).sum() + (o * np.roll(o, -1, 1)).sum())


def test_single_atom_random_walk():
    rows = cols = 5
    occ = np.zeros(25, np.int8)
    occ[12] = 1
    for n_ev in (1, 2, 17, 400):
        r = O.kmc_run(rows, cols, 0.1, 800.0, n_ev, occ, O.mt64_state(5, 0))
        assert r["attempts"] == n_ev
        assert r["final_time"] == n_ev / 4.0
        assert np.array_equal(r["engine"], O.mt64_state(5, 2 * n_ev))
        site, dx, dy, jumps = r["atoms"][0]
        assert jumps == n_ev and (abs(dx) + abs(dy)) % 2 == n_ev % 2
        assert site == ((12 % rows + dx) % rows) + rows * ((12 // rows + dy) % cols)
    # squared displacement over windows of m = 100 events (time grid at 25)
    r = O.kmc_run(rows, cols, 0.1, 800.0, 400000, occ, O.mt64_state(6, 0), time_period=25.0)
    # a sample at time 25k comes before the event that reaches it (event 100k)
    assert r["events"][0] == 0 and np.all(r["events"][1:] == 100 * np.arange(1, 4001) - 1)
    x, y = r["sum_dx"][1:].astype(float), r["sum_dy"][1:].astype(float)
    d2x, d2y = np.diff(x) ** 2, np.diff(y) ** 2
    for d2, want in ((d2x + d2y, 100.0), (d2x, 50.0), (d2y, 50.0)):
        se = d2.std(ddof=1) / np.sqrt(len(d2))
        assert abs(d2.mean() - want) < 4.5 * se, (d2.mean(), want, se)


def _canonical_average(rows, cols, n_atoms, J, T):
    n = rows * cols
    beta = 1.0 / (KB * T)
    Ss, ws = [], []
    for c in itertools.combinations(range(n), n_atoms):
        o = np.zeros(n, int)
        o[list(c)] = 1
        S = _bond_sum(o, rows, cols)
        Ss.append(S)
        ws.append(np.exp(beta * J * S))  # E = -J S
    ws = np.array(ws)
    return float(np.dot(Ss, ws) / ws.sum())


def test_time_averages_are_canonical():
    rows, cols, J, T = 3, 4, 0.05, 600.0
    occ = np.zeros(12, np.int8)
    occ[[0, 2, 5, 7, 11]] = 1
    exact = _canonical_average(rows, cols, 5, J, T)
    r = O.kmc_run(rows, cols, J, T, 1500000, occ, O.mt64_state(8, 0), time_period=1.0)
    S = r["S"].astype(float)
    nb = 100
    bm = S[: nb * (len(S) // nb)].reshape(nb, -1).mean(axis=1)
    se = bm.std(ddof=1) / np.sqrt(nb)
    assert abs(S.mean() - exact) < 4.5 * se, (S.mean(), exact, se)
    # averages over events weight states by their exit rates instead: far off
    p = O.kmc_run(rows, cols, J, T, 1500000, occ, O.mt64_state(8, 0))
    Sp = p["S"].astype(float)
    bp = Sp[: nb * (len(Sp) // nb)].reshape(nb, -1).mean(axis=1)
    sep = bp.std(ddof=1) / np.sqrt(nb)
    assert abs(Sp.mean() - exact) > 10 * max(se, sep)


def test_no_draw_without_interactions():
    rows, cols = 6, 5
    rng = np.random.default_rng(3)
    occ = (rng.random(30) < 0.4).astype(np.int8)
    r = O.kmc_run(rows, cols, 0.0, 900.0, 300, occ, O.mt64_state(7, 0))
    assert np.array_equal(r["engine"], O.mt64_state(7, 2 * r["attempts"]))
    assert r["attempts"] > 300  # some attempts hit occupied sites


def test_lists_and_atoms_consistent_after_each_event():
    rows, cols = 5, 6
    rng = np.random.default_rng(4)
    occ0 = (rng.random(30) < 0.5).astype(np.int8)
    n_atoms = int(occ0.sum())
    for n_ev in range(1, 41):
        r = O.kmc_run(rows, cols, 0.08, 700.0, n_ev, occ0, O.mt64_state(9, 0))
        occ, loc, atoms = r["occ"], r["loc"], r["atoms"]
        n0 = int((occ == 0).sum())
        assert sorted(loc[:n0]) == [l for l in range(30) if occ[l] == 0]
        assert sorted(loc[n0:]) == [l for l in range(30) if occ[l] == 1]
        assert int(occ.sum()) == n_atoms
        assert sorted(atoms[:, 0]) == [l for l in range(30) if occ[l] == 1]
        assert int(atoms[:, 3].sum()) == n_ev
        for (site, dx, dy, _), l0 in zip(atoms, [l for l in range(30) if occ0[l]]):
            assert site == ((l0 % rows + dx) % rows) + rows * ((l0 // rows + dy) % cols)
