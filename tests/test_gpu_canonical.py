"""CANONICAL mode on the GPU vs the oracle: bit-exact pair-swap trajectories.

The oracle (refcpu::canonical_run) runs the generic occupation_metropolis loop
(occupation_metropolis.hh:87-153) with propose_canonical_event
(OccEventProposal.hh:109-236) over an OccLocation (OccLocation.cc:39-118,
216-283) and std::mt19937_64; the kernel must reproduce, per pass, the FP64
running energy (bitwise) and the bond sum, and at the end the acceptance
counts, the occupation, both candidate lists (their order decides later
draws) and the engine state words.
"""
import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O

pytestmark = pytest.mark.gpu

J = 0.1


def _spins(rng, rows, cols, n_up):
    s = -np.ones(rows * cols, dtype=np.int8)
    s[rng.permutation(rows * cols)[:n_up]] = 1
    return s


def _compare(r, c, o, rows, cols):
    rec = r.samples(c)
    np.testing.assert_array_equal(rec["e_run"].view(np.uint64), o["e_run"].view(np.uint64))
    np.testing.assert_array_equal(rec["S1"], o["S"])
    n_pass, n_acc, n_rej = r.counters(c)
    assert (n_acc, n_rej) == (o["n_accept"], o["n_reject"])
    np.testing.assert_array_equal(r.get_occupation(c), o["occ"])
    np.testing.assert_array_equal(r.occ_location(c), o["loc"])
    np.testing.assert_array_equal(r.rng_dump(c), o["engine"])


@pytest.mark.parametrize("rows,cols,n_up,T,seed,passes,opts", [
    (8, 8, 32, 2000.0, 5, 60, 0),
    (25, 25, 300, 1500.0, 42, 60, 0),
    (16, 16, 1, 3000.0, 3, 40, 0),
    (16, 16, 255, 900.0, 4, 40, 0),
    (7, 13, 80, 800.0, 9, 60, 0),
    (2, 9, 9, 2500.0, 11, 80, 0),
    (1, 12, 5, 1000.0, 12, 80, 0),
    (64, 64, 1500, 2200.0, 13, 15, 0),
    (128, 128, 8192, 2000.0, 1000, 6, 0),
    (25, 25, 300, 1500.0, 42, 30, cm.OPT_EXACT_SCALAR_WALK),
])
def test_canonical_bit_exact(rows, cols, n_up, T, seed, passes, opts):
    print("case:", rows, cols, n_up, T, seed, passes, opts, flush=True)
    rng = np.random.default_rng(seed)
    occ = _spins(rng, rows, cols, n_up)
    o = O.canonical_run(rows, cols, J, T, passes, occ, O.mt64_state(seed, 0))
    r = cm.Run([dict(shape=(rows, cols), temperature=T, seed=seed)], mode=cm.MODE_CANONICAL,
               J1=J, options=opts)
    r.set_occupation(0, occ)
    for k in (1, passes // 3, passes - 1 - passes // 3):
        r.run(k)
    _compare(r, 0, o, rows, cols)
    assert r.counters(0)[0] == passes
    r.close()


def test_canonical_batch_and_fold():
    """8 chains of different shapes, temperatures and compositions in one launch;
    use_nlist=False samples the row/column fold of the formation energy."""
    rng = np.random.default_rng(77)
    chains, occs = [], []
    shapes = [(32, 32), (20, 30), (48, 16), (10, 10), (33, 31), (64, 40), (12, 50), (24, 24)]
    for i, (a, b) in enumerate(shapes):
        chains.append(dict(shape=(a, b), temperature=600.0 + 350.0 * i, seed=1000 + i))
        occs.append(_spins(rng, a, b, int((0.1 + 0.1 * i) * a * b)))
    r = cm.Run(chains, mode=cm.MODE_CANONICAL, J1=J, use_nlist=False)
    for i, o in enumerate(occs):
        r.set_occupation(i, o)
    r.run(25)
    for i, ((a, b), occ) in enumerate(zip(shapes, occs)):
        print("chain", i, a, b, flush=True)
        o = O.canonical_run(a, b, J, chains[i]["temperature"], 25, occ,
                            O.mt64_state(1000 + i, 0), use_nlist=False)
        _compare(r, i, o, a, b)
        e_fold = r.samples(i)["e_fold"][-1]
        ref = O.ising_energy(o["occ"].astype(np.int32), a, b, J, False)
        assert np.float64(e_fold).view(np.uint64) == np.float64(ref).view(np.uint64)
    r.close()


def test_canonical_lists_handover():
    """Loading the lists, occupation and engine of an oracle run and continuing
    equals the oracle's uninterrupted run."""
    rows, cols, seed, T = 20, 20, 8, 1800.0
    occ = _spins(np.random.default_rng(8), rows, cols, 170)
    full = O.canonical_run(rows, cols, J, T, 12, occ, O.mt64_state(seed, 0))
    head = O.canonical_run(rows, cols, J, T, 5, occ, O.mt64_state(seed, 0))
    r = cm.Run([dict(shape=(rows, cols), temperature=T, seed=seed)], mode=cm.MODE_CANONICAL,
               J1=J)
    r.set_occupation(0, head["occ"])
    r.set_occ_location(0, head["loc"])
    r.rng_load(0, head["engine"])
    r.run(7)
    np.testing.assert_array_equal(r.get_occupation(0), full["occ"])
    np.testing.assert_array_equal(r.occ_location(0), full["loc"])
    np.testing.assert_array_equal(r.rng_dump(0), full["engine"])
    np.testing.assert_array_equal(r.samples(0)["S1"], full["S"][5:])
    r.close()


def test_canonical_errors():
    r = cm.Run([dict(shape=(4, 4), temperature=1000.0, seed=1)], mode=cm.MODE_CANONICAL, J1=J)
    with pytest.raises(cm.McgError, match="No events possible"):
        r.run(1)  # all +1: no species-0 site
    occ = _spins(np.random.default_rng(1), 4, 4, 8)
    r.set_occupation(0, occ)
    bad = np.arange(16)
    with pytest.raises(cm.McgError, match="do not match"):
        r.set_occ_location(0, bad)
    with pytest.raises(cm.McgError, match="exactly once"):
        r.set_occ_location(0, np.zeros(16, dtype=np.int64))
    r.run(2)
    r.close()


def test_canonical_many_short_launches():
    """One-pass launches (each starts and ends the engine hand-off between the
    walking and the twisting wavefront) equal one uninterrupted run."""
    rows, cols, n_up, T, seed, passes = 6, 6, 14, 1800.0, 21, 50
    occ = _spins(np.random.default_rng(seed), rows, cols, n_up)
    o = O.canonical_run(rows, cols, J, T, passes, occ, O.mt64_state(seed, 0))
    r = cm.Run([dict(shape=(rows, cols), temperature=T, seed=seed)], mode=cm.MODE_CANONICAL,
               J1=J)
    r.set_occupation(0, occ)
    for _ in range(passes):
        r.run(1)
    _compare(r, 0, o, rows, cols)
    r.close()
