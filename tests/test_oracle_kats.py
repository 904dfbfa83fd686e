"""Pin the oracle against the reference's own known answers and against what the
mathematics fixes (CPU only).

Known answers taken from the reference's tests:
  tests/unit/monte/Ising_basic_semigrand_canonical_test.cpp:113-267 (energies,
  compositions and SGC potential of a 25x25 all +1 lattice, J=0.1, mu=2);
  python/tests/sampling/test_CompletionCheck.py:5-37 (max_count cutoff);
  the C++ standard's mt19937_64 known answer ([rand.predef]).
"""
import math

import numpy as np
import pytest

import oracle as O

J = 0.1


def colmajor(a):
    """2D numpy array (rows, cols) -> reference linear order l = i + rows*j."""
    return np.asarray(a, dtype=np.int32).reshape(-1, order="F")


# --- reference KATs (Ising_basic_semigrand_canonical_test.cpp) -------------------

@pytest.mark.parametrize("use_nlist", [True, False])
def test_formation_energy_kat(use_nlist):
    occ = np.ones(25 * 25, dtype=np.int32)
    E = O.ising_energy(occ, 25, 25, J, use_nlist)
    assert math.isclose(E, 25 * 25 * 2.0 * -J, rel_tol=1e-12)  # :138-142
    assert math.isclose(E / 625, 2.0 * -J, rel_tol=1e-12)      # :145-149
    # flipping site 0 to -1: dEf = 8J (:159); to the same value: 0
    assert O.ising_delta(occ, 25, 25, J, use_nlist, 0.0, 0, -1) == pytest.approx(8.0 * J)
    assert O.ising_delta(occ, 25, 25, J, use_nlist, 0.0, 0, 1) == 0.0


@pytest.mark.parametrize("use_nlist", [True, False])
def test_sgc_potential_kat(use_nlist):
    mu = 2.0
    occ = np.ones(625, dtype=np.int32)
    E = O.sgc_potential(occ, 25, 25, J, use_nlist, mu)
    assert math.isclose(E, 625 * (2.0 * -J - mu * 1.0), rel_tol=1e-12)  # :244-246
    dE = O.ising_delta(occ, 25, 25, J, use_nlist, mu, 0, -1)
    assert dE == pytest.approx(8.0 * J - mu * (-1))  # :259-261
    assert O.ising_delta(occ, 25, 25, J, use_nlist, mu, 0, 1) == 0.0


def test_completion_max_count_kat():
    """test_CompletionCheck.py:5-37: max_count=12 stops at count 12 -> 12 passes."""
    r = O.sgc_run(5, 5, max_count=12, seed=1)
    assert r["n_pass"] == 12
    assert r["n_samples"] == 12
    assert r["n_accept"] + r["n_reject"] == 12 * 25


def test_mt19937_64_kat():
    x = O.mt64_outputs(5489, 10000)
    assert int(x[-1]) == 9981545732273789042


# --- closed forms ----------------------------------------------------------------

@pytest.mark.parametrize("use_nlist", [True, False])
def test_checkerboard_antiferro_energy(use_nlist):
    L = 8
    a = np.fromfunction(lambda i, j: np.where((i + j) % 2 == 0, 1, -1), (L, L))
    E = O.ising_energy(colmajor(a), L, L, J, use_nlist)
    assert math.isclose(E / (L * L), +2.0 * J, rel_tol=1e-12)


@pytest.mark.parametrize("use_nlist", [True, False])
@pytest.mark.parametrize("shape", [(6, 6), (5, 9), (7, 4)])
def test_delta_equals_energy_difference(use_nlist, shape):
    """E(after) - E(before) == dE for random flips on random lattices (brute force)."""
    rows, cols = shape
    rng = np.random.default_rng(3)
    for trial in range(40):
        occ = rng.choice([-1, 1], size=rows * cols).astype(np.int32)
        mu = float(rng.uniform(-1, 1))
        l = int(rng.integers(rows * cols))
        before = O.sgc_potential(occ, rows, cols, J, use_nlist, mu)
        after_occ = occ.copy()
        after_occ[l] = -occ[l]
        after = O.sgc_potential(after_occ, rows, cols, J, use_nlist, mu)
        dE = O.ising_delta(occ, rows, cols, J, use_nlist, mu, l, int(-occ[l]))
        assert dE == pytest.approx(after - before, abs=1e-9)


def test_energy_variants_agree_to_rounding():
    rng = np.random.default_rng(5)
    occ = rng.choice([-1, 1], size=13 * 11).astype(np.int32)
    a = O.ising_energy(occ, 13, 11, J, True)
    b = O.ising_energy(occ, 13, 11, J, False)
    assert a == pytest.approx(b, rel=1e-12)


# --- RNG distribution restatement vs the library ----------------------------------

def test_canonical_clamp_and_rounding():
    assert O.canonical(0) == 0.0
    assert O.canonical(2**64 - 1) == np.nextafter(1.0, 0.0)  # rounds to 2^64 -> clamped
    assert O.canonical(2**63) == 0.5
    # round-to-nearest of the 64-bit word (ties to even at 2^53 granularity)
    x = (1 << 63) + (1 << 10)  # exactly half an ulp above 2^63 -> ties to even (2^63)
    assert O.canonical(x) == 0.5


def test_lemire_rejection_path():
    """uniform_int_distribution redraws while low < (2^64 - range) % range."""
    rng = (1 << 63) + 1  # threshold = (2^64 - range) % range = 2^63 - 1
    maximum = rng - 1
    # x = 1: low = range (not < range): accepted directly, value = hi = 0
    v, used = O.uniform_int_replay([1], maximum)
    assert used == 1 and v == 0
    # x = 2: low = 2*range mod 2^64 = 2 < threshold -> rejected; next word 1 accepted
    v, used = O.uniform_int_replay([2, 1], maximum)
    assert used == 2 and v == 0


def test_metropolis_probability_is_exp():
    kb = O.kb()
    assert kb == 8.6173303e-05
    for dE in [0.0, 0.2, 0.4, 0.8, 1.6]:
        assert O.accept_prob(dE, 2000.0) == pytest.approx(math.exp(-dE / (kb * 2000.0)),
                                                          rel=1e-15)


# --- loop invariants ---------------------------------------------------------------

def test_loop_invariants():
    r = O.sgc_run(6, 7, max_count=50, seed=9, sample_period=3, mu=0.3, temperature=1500.0)
    assert r["n_pass"] == 50
    assert r["n_accept"] + r["n_reject"] == 50 * 42
    assert r["n_samples"] == 50 // 3
    # sampled values are consistent with the integer observables
    N = 42
    x = (N + r["M"]) / 2.0 / N
    np.testing.assert_array_equal(r["param_composition"], x)
    ef = (r["S"].astype(float) * -J) / N
    np.testing.assert_array_equal(r["formation_energy"], ef)
