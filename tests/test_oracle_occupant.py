"""Pins of the oracle's occupant-event restatement (refcpu::occ_run): the
generic occupation_metropolis loop over an OccLocation with several asymmetric
units, with semi-grand canonical single-site events or canonical pair swaps
(occupation_metropolis.hh:87-153, OccEventProposal.hh:109-352,
OccLocation.cc:39-118, 216-283).

Independent checks, none of which retypes the oracle's code:
- the stationary distribution of the chain the reference's algorithm defines
  (swap type with probability weight / total, a uniform site of each chosen
  list, Metropolis acceptance), built as a transition matrix over every state
  of a small lattice and solved exactly; the oracle's long-run energy agrees
  within 4.5 sigma, and the same test rejects a plausible mistake (swap types
  chosen uniformly) by a wide margin;
- with one binary unit, no exchange potentials and the single canonical swap,
  the loop is the canonical loop already pinned (refcpu::canonical_run), bit
  for bit;
- the words a step draws: with J = 0 and no exchange potentials every delta is
  0, so every step draws the swap-type word, its Lemire words and one
  acceptance word, and the engine ends exactly that many outputs on;
- the OccLocation invariants the reference's own test checks after every
  apply (tests/unit/monte/OccLocation_test.cpp:15-65: every mutating site in
  the list of its unit and species, once), after each of the first steps;
- the "No events possible" stop of OccEventProposal.hh:136-139.
"""
import itertools

import numpy as np
import pytest

import oracle as O

KB = 8.6173303e-05

# a 2x4 lattice: unit 0 (A, B) on five sites, unit 1 (B, A) on two, unit 2
# (A only) on one
ROWS, COLS = 2, 4
SITE_ASYM = [0, 0, 0, 1, 0, 1, 2, 0]
SPEC = [[0, 1], [1, 0], [0]]
OCC0 = np.array([0, 1, 0, 0, 1, 1, 0, 0], dtype=np.int8)


def _energy(st, J, mu):
    s = [2 * o - 1 for o in st]
    b = 0
    for l in range(ROWS * COLS):
        i, j = l % ROWS, l // ROWS
        b += s[l] * (s[(i + 1) % ROWS + ROWS * j] + s[i + ROWS * ((j + 1) % COLS)])
    return -J * b - sum(mu[SPEC[SITE_ASYM[l]][st[l]]] for l in range(ROWS * COLS))


def _species_totals(st):
    return tuple(sum(1 for l in range(len(st)) if SPEC[SITE_ASYM[l]][st[l]] == s)
                 for s in range(2))


def _stationary(J, T, mu, swaps, canonical, uniform_types=False):
    """Stationary distribution of the reference's chain over every reachable
    state, from its transition matrix."""
    n = ROWS * COLS
    cands = O.occ_candidates(SPEC)
    mut = [l for l in range(n) if len(SPEC[SITE_ASYM[l]]) > 1]
    states = []
    for bits in itertools.product([0, 1], repeat=len(mut)):
        st = [0] * n
        for l, b in zip(mut, bits):
            st[l] = b
        if canonical and _species_totals(st) != _species_totals(OCC0):
            continue
        states.append(tuple(st))
    index = {s: i for i, s in enumerate(states)}
    E = np.array([_energy(s, J, mu) for s in states])
    beta = 1.0 / (KB * T)
    P = np.zeros((len(states), len(states)))
    for si, st in enumerate(states):
        lists = [[l for l in mut if SITE_ASYM[l] == b and SPEC[b][st[l]] == sp] for b, sp in cands]
        w = np.array([len(lists[a]) * len(lists[b]) if canonical else len(lists[a])
                      for a, b in swaps], dtype=float)
        if uniform_types:
            w = (w > 0).astype(float)
        for i, (ca, cb) in enumerate(swaps):
            if w[i] == 0:
                continue
            if canonical:
                picks = [(a, b) for a in lists[ca] for b in lists[cb]]
            else:
                picks = [(a,) for a in lists[ca]]
            for pk in picks:
                p = w[i] / w.sum() / len(picks)
                new = list(st)
                for l in pk:
                    new[l] = 1 - new[l]
                sj = index[tuple(new)]
                dE = E[sj] - E[si]
                acc = 1.0 if dE < 0 else np.exp(-beta * dE)
                P[si, sj] += p * acc
                P[si, si] += p * (1.0 - acc)
    vals, vecs = np.linalg.eig(P.T)
    k = int(np.argmin(abs(vals - 1.0)))
    pi = np.real(vecs[:, k])
    return E, pi / pi.sum()


@pytest.mark.parametrize("canonical", [False, True])
@pytest.mark.parametrize("J,T,mu", [(0.02, 3000.0, [0.0, 0.1]), (0.01, 2000.0, [0.0, 0.2])])
def test_stationary_distribution(canonical, J, T, mu):
    swaps = O.occ_canonical_swaps(SPEC) if canonical else O.occ_semigrand_swaps(SPEC)
    E, pi = _stationary(J, T, mu, swaps, canonical)
    _, pi_wrong = _stationary(J, T, mu, swaps, canonical, uniform_types=True)
    n_uc = ROWS * COLS
    exact, wrong = float(pi @ E) / n_uc, float(pi_wrong @ E) / n_uc
    r = O.occ_run(ROWS, COLS, SITE_ASYM, SPEC, mu, swaps, canonical, J, T, 7 * 300000, OCC0,
                  O.mt64_state(9, 0))
    assert r["error"] is None and len(r["e_run"]) == 300000
    e = r["e_run"]
    nb = 100
    bm = e[: nb * (len(e) // nb)].reshape(nb, -1).mean(axis=1)
    se = bm.std(ddof=1) / np.sqrt(nb)
    assert abs(e.mean() - exact) < 4.5 * se, (e.mean(), exact, se)
    # the mistake the test is there to catch is far outside that band
    assert abs(wrong - exact) > 20 * se


def test_one_binary_unit_is_the_canonical_loop():
    rows, cols = 9, 7
    n = rows * cols
    rng = np.random.default_rng(5)
    occ = rng.integers(0, 2, n).astype(np.int8)
    spins = (2 * occ - 1).astype(np.int8)
    spec = [[0, 1]]
    swaps = O.occ_canonical_swaps(spec)
    assert swaps == [(0, 1)]
    for T in (900.0, 2500.0):
        eng = O.mt64_state(21, 0)
        a = O.occ_run(rows, cols, np.zeros(n, np.int8), spec, [0.0, 0.0], swaps, True, 0.1, T,
                      40 * n, occ, eng)
        b = O.canonical_run(rows, cols, 0.1, T, 40, spins, eng)
        assert np.array_equal(a["e_run"].view(np.uint64), b["e_run"].view(np.uint64))
        assert np.array_equal(a["S"], b["S"])
        assert np.array_equal(2 * a["occ"].astype(int) - 1, b["occ"].astype(int))
        assert np.array_equal(a["loc"], b["loc"])
        assert np.array_equal(a["engine"], b["engine"])
        assert (a["n_accept"], a["n_reject"]) == (b["n_accept"], b["n_reject"])
        assert a["initial"] == b["initial"]


@pytest.mark.parametrize("canonical,words", [(False, 3), (True, 4)])
def test_words_per_step(canonical, words):
    swaps = O.occ_canonical_swaps(SPEC) if canonical else O.occ_semigrand_swaps(SPEC)
    for steps in (1, 2, 7, 500):
        r = O.occ_run(ROWS, COLS, SITE_ASYM, SPEC, [0.0, 0.0], swaps, canonical, 0.0, 1000.0,
                      steps, OCC0, O.mt64_state(3, 0))
        assert r["n_accept"] == steps  # dE = 0: exp(0) = 1 accepts every word
        assert np.array_equal(r["engine"], O.mt64_state(3, words * steps))


def _lists_consistent(occ, loc, spec, site_asym):
    cands = O.occ_candidates(spec)
    want = [[l for l in range(len(occ)) if site_asym[l] == b and spec[b][occ[l]] == sp]
            for b, sp in cands]
    k = 0
    for w in want:
        got = list(loc[k:k + len(w)])
        k += len(w)
        if sorted(got) != w:
            return False
    return k == len(loc)


@pytest.mark.parametrize("canonical", [False, True])
def test_occ_location_invariants_after_each_step(canonical):
    rows, cols = 6, 5
    n = rows * cols
    rng = np.random.default_rng(8)
    site_asym = rng.integers(0, 3, n).astype(np.int8)
    spec = [[0, 1], [1, 0], [0]]
    occ = np.array([0 if spec[b] == [0] else rng.integers(0, 2) for b in site_asym], np.int8)
    swaps = O.occ_canonical_swaps(spec) if canonical else O.occ_semigrand_swaps(spec)
    for steps in range(1, 61):
        r = O.occ_run(rows, cols, site_asym, spec, [0.0, 0.05], swaps, canonical, 0.02, 5000.0,
                      steps, occ, O.mt64_state(11, 0))
        assert r["n_steps"] == steps
        assert _lists_consistent(r["occ"], r["loc_sites"], spec, site_asym)
        fixed = [l for l in range(n) if len(spec[site_asym[l]]) == 1]
        assert all(r["occ"][l] == 0 for l in fixed)
        if canonical:
            before = [sum(1 for l in range(n) if spec[site_asym[l]][occ[l]] == s) for s in (0, 1)]
            after = [sum(1 for l in range(n) if spec[site_asym[l]][r["occ"][l]] == s)
                     for s in (0, 1)]
            assert before == after


def test_no_events_stop():
    # only the cross-unit swap ((0, A), (1, B)): every accepted swap removes one
    # of each, so the proposals run out of events
    cands = O.occ_candidates(SPEC)
    swaps = [(cands.index((0, 0)), cands.index((1, 1)))]
    r = O.occ_run(ROWS, COLS, SITE_ASYM, SPEC, [0.0, 0.0], swaps, True, 0.0, 1000.0, 1000,
                  OCC0, O.mt64_state(4, 0))
    assert r["error"] == "Error in choose_canonical_swap: No events possible."
    n_a = sum(1 for l in range(8) if SITE_ASYM[l] == 0 and OCC0[l] == 0)
    n_b = sum(1 for l in range(8) if SITE_ASYM[l] == 1 and SPEC[1][OCC0[l]] == 1)
    # J = 0: every swap is accepted, one word each for type, a, b and acceptance
    assert r["n_steps"] == r["n_accept"] == min(n_a, n_b)
    assert np.array_equal(r["engine"], O.mt64_state(4, 4 * min(n_a, n_b)))
