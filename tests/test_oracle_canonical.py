"""Pins of the oracle's canonical pair-swap chain (refcpu::canonical_run).

The reference's own tests of this path (tests/unit/monte/OccLocation_test.cpp)
check OccLocation's invariants after a million proposed-and-applied canonical
events; they hold no golden vectors.  The oracle is pinned here by
  * those invariants (every site listed exactly once, in the list of its
    species, at the position its Mol.loc names) after runs,
  * conservation of the composition,
  * the running FP64 energy against the energy recomputed from the final
    configuration,
  * the stream consumption the draw order fixes: each step takes one word for
    the swap-type choice, one per chosen Mol and one more only when dE >= 0, so
    3*steps + n_reject <= words used <= 4*steps,
  * the canonical distribution by exact enumeration of a 4x4 torus at fixed
    composition, and
  * a zero-temperature limit from a ground state (no accepted move raises E).
The order of the lists after an apply (which Mol moves into the freed slot)
follows OccLocation.cc:254-283 by reading; it changes later draws but none of
the checks above.
"""
import numpy as np
import pytest

import oracle

J = 0.1


def _spins(rng, rows, cols, n_up):
    s = -np.ones(rows * cols, dtype=np.int8)
    s[rng.permutation(rows * cols)[:n_up]] = 1
    return s


def _bond_sum(s, rows, cols):
    g = s.reshape(cols, rows).T.astype(np.int64)  # g[i, j] = s[i + rows*j]
    return int((g * np.roll(g, -1, 0)).sum() + (g * np.roll(g, -1, 1)).sum())


def _check_lists(r, occ, n):
    n0 = int((occ < 0).sum())
    loc, ml = r["loc"], r["mol_loc"]
    assert sorted(loc.tolist()) == list(range(n))
    for pos, l in enumerate(loc):
        species = 0 if pos < n0 else 1
        assert (occ[l] > 0) == (species == 1)
        assert ml[l] == (pos if species == 0 else pos - n0)


@pytest.mark.parametrize("rows,cols,n_up,T", [(8, 8, 32, 2000.0), (6, 10, 13, 1500.0),
                                              (5, 5, 1, 3000.0), (3, 7, 20, 800.0)])
def test_invariants_and_energy(rows, cols, n_up, T):
    rng = np.random.default_rng(rows * 100 + n_up)
    occ = _spins(rng, rows, cols, n_up)
    n = rows * cols
    r = oracle.canonical_run(rows, cols, J, T, 20, occ, oracle.mt64_state(7, 0))
    assert int((r["occ"] > 0).sum()) == n_up
    _check_lists(r, r["occ"], n)
    # running energy vs recomputed (to rounding of the running sum)
    e_final = -J * _bond_sum(r["occ"], rows, cols) / n
    assert abs(r["e_run"][-1] - e_final) < 1e-12
    assert r["S"][-1] == _bond_sum(r["occ"], rows, cols)
    assert abs(r["initial"] - (-J * _bond_sum(occ, rows, cols) / n)) < 1e-15
    assert r["n_accept"] + r["n_reject"] == 20 * n


def test_fold_variant_initial_energy():
    """The row/column fold and the neighbour-list sum give the same dynamics;
    only the initial per_supercell value may differ in its last bits."""
    occ = _spins(np.random.default_rng(4), 12, 10, 50)
    a = oracle.canonical_run(12, 10, J, 1500.0, 4, occ, oracle.mt64_state(3, 0))
    b = oracle.canonical_run(12, 10, J, 1500.0, 4, occ, oracle.mt64_state(3, 0), use_nlist=False)
    assert a["initial"] == pytest.approx(b["initial"], rel=1e-14, abs=1e-15)
    assert b["initial"] == oracle.ising_energy(occ.astype(np.int32), 12, 10, J, False) / 120
    np.testing.assert_array_equal(a["S"], b["S"])
    np.testing.assert_array_equal(a["loc"], b["loc"])


def test_lists_resume():
    """Running 2 + 3 passes with the lists handed over equals running 5."""
    rng = np.random.default_rng(3)
    occ = _spins(rng, 8, 8, 24)
    a = oracle.canonical_run(8, 8, J, 2500.0, 5, occ, oracle.mt64_state(11, 0))
    b = oracle.canonical_run(8, 8, J, 2500.0, 2, occ, oracle.mt64_state(11, 0))
    c = oracle.canonical_run(8, 8, J, 2500.0, 3, b["occ"], b["engine"], loc=b["loc"])
    assert np.array_equal(a["occ"], c["occ"])
    assert np.array_equal(a["loc"], c["loc"])
    assert np.array_equal(a["engine"], c["engine"])
    np.testing.assert_array_equal(a["S"][2:], c["S"])


def test_stream_consumption():
    rows = cols = 8
    rng = np.random.default_rng(5)
    occ = _spins(rng, rows, cols, 30)
    seed = 99
    r = oracle.canonical_run(rows, cols, J, 2000.0, 1, occ, oracle.mt64_state(seed, 0))
    steps = rows * cols
    lo, hi = 3 * steps + r["n_reject"], 4 * steps
    used = [k for k in range(lo, hi + 4)
            if np.array_equal(oracle.mt64_state(seed, k), r["engine"])]
    assert len(used) == 1 and lo <= used[0] <= hi


def _enumerate(rows, cols, n_up, T):
    n = rows * cols
    beta = 1.0 / (oracle.kb() * T)
    es = []
    for m in range(1 << n):
        if bin(m).count("1") != n_up:
            continue
        s = np.array([1 if (m >> l) & 1 else -1 for l in range(n)], dtype=np.int8)
        es.append(-J * _bond_sum(s, rows, cols) / n)
    es = np.array(es)
    w = np.exp(-beta * (es - es.min()) * n)
    return float((w * es).sum() / w.sum())


@pytest.mark.parametrize("n_up,T", [(8, 2000.0), (5, 1200.0)])
def test_canonical_distribution(n_up, T):
    rows = cols = 4
    exact = _enumerate(rows, cols, n_up, T)
    rng = np.random.default_rng(n_up)
    occ = _spins(rng, rows, cols, n_up)
    r = oracle.canonical_run(rows, cols, J, T, 40000, occ, oracle.mt64_state(2024 + n_up, 0))
    e = r["e_run"][1000:]
    blocks = e[: len(e) // 100 * 100].reshape(100, -1).mean(axis=1)
    err = blocks.std(ddof=1) / np.sqrt(len(blocks))
    assert abs(blocks.mean() - exact) < 4 * err + 1e-12, (blocks.mean(), exact, err)


def test_zero_temperature_ground_state():
    """Columns 0-3 up, 4-7 down on an 8x8 torus is a ground state at M = 0; at 1 K
    no move that raises the energy is accepted, so E stays at its minimum."""
    rows = cols = 8
    g = np.ones((rows, cols), dtype=np.int8)
    g[:, 4:] = -1
    occ = g.T.reshape(-1).copy()  # column-major
    r = oracle.canonical_run(rows, cols, J, 1.0, 10, occ, oracle.mt64_state(1, 0))
    assert np.all(r["e_run"] == r["initial"])
    assert np.all(r["S"] == _bond_sum(occ, rows, cols))


def test_no_events_possible():
    occ = np.ones(16, dtype=np.int8)
    with pytest.raises(oracle.CanonicalError, match="No events possible"):
        oracle.canonical_run(4, 4, J, 1000.0, 1, occ, oracle.mt64_state(1, 0))
