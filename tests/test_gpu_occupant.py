"""OCCUPANT mode on the GPU against the oracle (refcpu::occ_run): the generic
occupation_metropolis loop over an OccLocation with several asymmetric units,
semi-grand canonical single-site events and canonical pair swaps
(occupation_metropolis.hh:87-153, OccEventProposal.hh:109-352,
OccLocation.cc:39-118, 216-283).  Bit-exact: the FP64 running energy per pass,
bond sum and spin sum per pass, final occupation, both kinds of lists in order,
engine words and acceptance counts."""
import numpy as np
import pytest

import casmcode_monte_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

SPEC3 = [[0, 1], [1, 0], [0]]  # units (A, B), (B, A), (A)


def _units(rows, cols, seed, n_units=3):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n_units, rows * cols).astype(np.int8)


def _occ(site_asym, spec, seed, frac=0.5):
    rng = np.random.default_rng(seed + 1000)
    return np.array([0 if len(spec[b]) == 1 else int(rng.random() < frac) for b in site_asym],
                    dtype=np.int8)


def _handle(chains, spec, swaps, canonical, J=0.1, use_nlist=True, mus=None, n_species=2):
    occd = dict(asym_species=spec, n_species=n_species, canonical=canonical, swaps=swaps)
    if mus is not None:
        occd["exchange_potential"] = mus
    return M.Run(chains, mode=M.MODE_OCCUPANT, J1=J, use_nlist=use_nlist, sample_period=1,
                 occupant=occd)


def _gpu_state(r, c):
    s = r.samples(c)
    return dict(e_run=s["e_run"], S=s["S1"], M=s["M"], occ=r.get_occupation(c),
                loc=r.occ_location(c), engine=r.rng_dump(c), counters=r.counters(c),
                position=r.position(c))


def _assert_equal(g, o, tag):
    n = len(o["e_run"])
    assert len(g["e_run"]) == n, tag
    assert np.array_equal(g["e_run"].view(np.uint64), o["e_run"].view(np.uint64)), tag
    assert np.array_equal(g["S"], o["S"]), tag
    assert np.array_equal(g["M"], o["M"]), tag
    assert np.array_equal(g["occ"], o["occ"]), tag
    assert np.array_equal(g["loc"], o["loc_sites"]), tag
    assert np.array_equal(g["engine"], o["engine"]), tag
    assert g["counters"][1:] == (o["n_accept"], o["n_reject"]), tag


def _case(rows, cols, site_asym, spec, mu, canonical, T, passes, seed, J=0.1, use_nlist=True,
          swaps=None, occ=None):
    if swaps is None:
        swaps = O.occ_canonical_swaps(spec) if canonical else O.occ_semigrand_swaps(spec)
    if occ is None:
        occ = _occ(site_asym, spec, seed)
    n_mut = sum(1 for b in site_asym if len(spec[b]) > 1)
    print(f"occupant case {rows}x{cols} canonical={canonical} T={T} seed={seed}", flush=True)
    with _handle([dict(shape=(rows, cols), temperature=T, seed=seed)], spec, swaps, canonical,
                 J=J, use_nlist=use_nlist, mus=[mu], n_species=len(mu)) as r:
        r.set_site_asym(0, site_asym)
        r.set_occupation(0, occ)
        assert np.array_equal(r.rng_dump(0), O.mt64_state(seed, 0))
        r.run(passes)
        g = _gpu_state(r, 0)
    o = O.occ_run(rows, cols, site_asym, spec, mu, swaps, canonical, J, T, passes * n_mut, occ,
                  O.mt64_state(seed, 0), use_nlist=use_nlist)
    assert o["error"] is None
    _assert_equal(g, o, f"{rows}x{cols} canonical={canonical} T={T}")
    assert g["position"][0] == passes and g["position"][1] == 0
    return g, o


@pytest.mark.parametrize("T", [600.0, 1500.0, 4000.0, 1e5])
def test_semigrand_one_unit(T):
    _case(16, 16, np.zeros(256, np.int8), [[0, 1]], [0.0, 0.05], False, T, 30, seed=3)


@pytest.mark.parametrize("T", [800.0, 2500.0])
def test_semigrand_three_units(T):
    ua = _units(24, 20, 4)
    _case(24, 20, ua, SPEC3, [0.0, -0.03], False, T, 25, seed=5)


@pytest.mark.parametrize("T", [900.0, 3000.0])
def test_canonical_three_units(T):
    ua = _units(16, 18, 6)
    _case(16, 18, ua, SPEC3, [0.0, 0.02], True, T, 25, seed=7)


def test_semigrand_species_order_and_exchange_potentials():
    # four units, three species: (A, B), (C, A), (B, C), (B)
    spec = [[0, 1], [2, 0], [1, 2], [1]]
    ua = _units(20, 14, 8, n_units=4)
    _case(20, 14, ua, spec, [0.01, -0.04, 0.07], False, 1800.0, 20, seed=9)


def test_fold_variant_initial_energy():
    ua = _units(12, 10, 10)
    _case(12, 10, ua, SPEC3, [0.0, 0.05], False, 2000.0, 10, seed=11, use_nlist=False)


@pytest.mark.parametrize("rows,cols", [(1, 7), (7, 1), (2, 5), (5, 2), (2, 2), (1, 1)])
@pytest.mark.parametrize("canonical", [False, True])
def test_thin_lattices(rows, cols, canonical):
    n = rows * cols
    spec = [[0, 1]]
    occ = np.array([l % 2 for l in range(n)], np.int8)
    if canonical and n == 1:
        pytest.skip("one site has no pair to swap")
    _case(rows, cols, np.zeros(n, np.int8), spec, [0.0, 0.1], canonical, 2500.0, 40, seed=12,
          occ=occ)


def test_one_unit_canonical_equals_canonical_mode():
    rows, cols = 18, 14
    n = rows * cols
    occ = _occ(np.zeros(n, np.int8), [[0, 1]], 13)
    spins = (2 * occ.astype(int) - 1).astype(np.int8)
    g, _ = _case(rows, cols, np.zeros(n, np.int8), [[0, 1]], [0.0, 0.0], True, 1700.0, 30,
                 seed=13, occ=occ)
    with M.Run([dict(shape=(rows, cols), temperature=1700.0, seed=13)], mode=M.MODE_CANONICAL,
               J1=0.1, sample_period=1) as r:
        r.set_occupation(0, spins)
        r.run(30)
        s = r.samples(0)
        assert np.array_equal(s["e_run"].view(np.uint64), g["e_run"].view(np.uint64))
        assert np.array_equal(r.get_occupation(0), 2 * g["occ"].astype(np.int8) - 1)
        assert np.array_equal(r.occ_location(0), g["loc"])
        assert np.array_equal(r.rng_dump(0), g["engine"])


@pytest.mark.parametrize("canonical", [False, True])
def test_run_steps_chunks(canonical):
    rows, cols = 10, 9
    ua = _units(rows, cols, 14)
    spec = SPEC3
    swaps = O.occ_canonical_swaps(spec) if canonical else O.occ_semigrand_swaps(spec)
    occ = _occ(ua, spec, 14)
    n_mut = sum(1 for b in ua if len(spec[b]) > 1)
    mu = [0.0, 0.04]
    with _handle([dict(shape=(rows, cols), temperature=2200.0, seed=15)], spec, swaps, canonical,
                 mus=[mu]) as r:
        r.set_site_asym(0, ua)
        r.set_occupation(0, occ)
        total = 0
        for k in [1, 2, 37, n_mut - 40, 3 * n_mut + 5, 0, 11]:
            r.run_steps(k)
            total += k
            o = O.occ_run(rows, cols, ua, spec, mu, swaps, canonical, 0.1, 2200.0, total, occ,
                          O.mt64_state(15, 0))
            g = _gpu_state(r, 0)
            _assert_equal(g, o, f"after {total} steps")
            assert g["position"][:2] == (total // n_mut, total % n_mut)
        r.run(2)
        total += 2 * n_mut
        o = O.occ_run(rows, cols, ua, spec, mu, swaps, canonical, 0.1, 2200.0, total, occ,
                      O.mt64_state(15, 0))
        _assert_equal(_gpu_state(r, 0), o, "passes after steps")


def test_batch_of_chains_equals_each_alone():
    shapes = [(8, 8), (16, 12), (5, 9), (20, 20), (3, 4)]
    Ts = [700.0, 1500.0, 2600.0, 1200.0, 5000.0]
    mus = [[0.0, 0.01 * k] for k in range(len(shapes))]
    spec = SPEC3
    swaps = O.occ_semigrand_swaps(spec)
    chains = [dict(shape=s, temperature=T, seed=20 + k) for k, (s, T) in enumerate(zip(shapes, Ts))]
    units = [_units(s[0], s[1], 30 + k) for k, s in enumerate(shapes)]
    occs = [_occ(u, spec, 40 + k) for k, u in enumerate(units)]
    print("occupant batch of 5 chains", flush=True)
    with _handle(chains, spec, swaps, False, mus=mus) as r:
        for k in range(len(shapes)):
            r.set_site_asym(k, units[k])
            r.set_occupation(k, occs[k])
        r.run(12)
        for k, s in enumerate(shapes):
            n_mut = sum(1 for b in units[k] if len(spec[b]) > 1)
            o = O.occ_run(s[0], s[1], units[k], spec, mus[k], swaps, False, 0.1, Ts[k], 12 * n_mut,
                          occs[k], O.mt64_state(20 + k, 0))
            _assert_equal(_gpu_state(r, k), o, f"chain {k}")


def test_set_occ_location_custom_order():
    rows, cols = 9, 11
    ua = _units(rows, cols, 16)
    spec = SPEC3
    swaps = O.occ_semigrand_swaps(spec)
    occ = _occ(ua, spec, 16)
    with _handle([dict(shape=(rows, cols), temperature=1900.0, seed=17)], spec, swaps, False,
                 mus=[[0.0, 0.0]]) as r:
        r.set_site_asym(0, ua)
        r.set_occupation(0, occ)
        loc = r.occ_location(0)
        # reverse every candidate list
        mol_site = [l for l in range(rows * cols) if len(spec[ua[l]]) > 1]
        site_mol = {l: i for i, l in enumerate(mol_site)}
        cands = O.occ_candidates(spec)
        sizes = [sum(1 for l in range(rows * cols) if ua[l] == b and spec[b][occ[l]] == sp)
                 for b, sp in cands]
        new, k = [], 0
        for sz in sizes:
            new += list(loc[k:k + sz])[::-1]
            k += sz
        r.set_occ_location(0, new)
        assert list(r.occ_location(0)) == new
        r.run(8)
        g = _gpu_state(r, 0)
    o = O.occ_run(rows, cols, ua, spec, [0.0, 0.0], swaps, False, 0.1, 1900.0, 8 * len(mol_site),
                  occ, O.mt64_state(17, 0), loc=[site_mol[l] for l in new])
    _assert_equal(g, o, "custom list order")


def test_no_events_stop_matches_the_reference_loop():
    rows, cols = 6, 6
    ua = _units(rows, cols, 18)
    spec = SPEC3
    cands = O.occ_candidates(spec)
    swaps = [(cands.index((0, 0)), cands.index((1, 1)))]
    occ = _occ(ua, spec, 18)
    with _handle([dict(shape=(rows, cols), temperature=50000.0, seed=19)], spec, swaps, True,
                 J=0.0, mus=[[0.0, 0.0]]) as r:
        r.set_site_asym(0, ua)
        r.set_occupation(0, occ)
        with pytest.raises(M.McgError, match="No events possible"):
            r.run(100)
        g = _gpu_state(r, 0)
        with pytest.raises(M.McgError, match="No events possible"):
            r.run(1)
    n_mut = sum(1 for b in ua if len(spec[b]) > 1)
    o = O.occ_run(rows, cols, ua, spec, [0.0, 0.0], swaps, True, 0.0, 50000.0, 100 * n_mut, occ,
                  O.mt64_state(19, 0))
    assert o["error"] == "Error in choose_canonical_swap: No events possible."
    _assert_equal(g, o, "stopped run")
    p, st, _ = g["position"]
    assert p * n_mut + st == o["n_steps"]


def test_largest_lattice():
    rows, cols = 200, 200
    ua = _units(rows, cols, 21, n_units=2)
    _case(rows, cols, ua, [[0, 1], [1, 0]], [0.0, 0.03], False, 2000.0, 3, seed=22)


def test_invalid_input_is_refused():
    spec = SPEC3
    ua = _units(6, 6, 23)
    with _handle([dict(shape=(6, 6), temperature=1000.0, seed=1)], spec,
                 O.occ_semigrand_swaps(spec), False) as r:
        r.set_site_asym(0, ua)
        bad = _occ(ua, spec, 23)
        fixed = [l for l in range(36) if ua[l] == 2]
        bad[fixed[0]] = 1
        with pytest.raises(M.McgError, match="occupant indices"):
            r.set_occupation(0, bad)
        with pytest.raises(M.McgError, match="out of range"):
            r.set_site_asym(0, np.full(36, 3, np.int8))
    with pytest.raises(M.McgError, match="semi-grand"):
        _handle([dict(shape=(6, 6), temperature=1000.0)], spec, [(0, 2)], False)
    with pytest.raises(M.McgError, match="canonical swap"):
        _handle([dict(shape=(6, 6), temperature=1000.0)], spec, [(0, 1), (0, 3)], True)
