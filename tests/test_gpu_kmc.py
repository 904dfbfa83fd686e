"""KMC mode on the GPU against the oracle (refcpu::kmc_run): the loop of
kinetic_monte_carlo.hh:136-229 with OccLocation atom tracking for a square
lattice gas under the rejection event selector.  Bit-exact: records (pass or
event count, time, bond sum, squared-displacement and displacement sums),
final occupation, both lists in order, every atom's site, displacement and
jumps, the loop's time, the engine words and the attempt counts."""
import numpy as np
import pytest

import casmcode_monte_amd as M
import oracle as O

pytestmark = pytest.mark.gpu


def _occ(n, frac, seed):
    return (np.random.default_rng(seed).random(n) < frac).astype(np.int8)


def _gpu_state(r, c, n_atoms):
    s = r.samples(c)
    t, site, _ = r.nfold_state(c)
    return dict(rec=s, occ=r.get_occupation(c), loc=r.occ_location(c), atoms=r.atoms(c, n_atoms),
                engine=r.rng_dump(c), counters=r.counters(c), position=r.position(c), time=t,
                pending=site)


def _assert_equal(g, o, period, tag):
    rec = g["rec"]
    assert len(rec) == o["n_records"], tag
    if period > 0:
        assert np.array_equal(rec["pass"], o["events"]), tag
    assert np.array_equal(rec["time"].view(np.uint64), o["time"].view(np.uint64)), tag
    assert np.array_equal(rec["S1"], o["S"]), tag
    assert np.array_equal(rec["S2"], o["sum_sq"]), tag
    assert np.array_equal(rec["S3"], o["sum_dx"]), tag
    assert np.array_equal(rec["T3"], o["sum_dy"]), tag
    assert np.array_equal(g["occ"], o["occ"]), tag
    assert np.array_equal(g["loc"], o["loc"]), tag
    assert np.array_equal(g["atoms"], o["atoms"]), tag
    assert np.array_equal(g["engine"], o["engine"]), tag
    assert g["time"] == o["final_time"] and g["pending"] == -1, tag
    n_pass, n_acc, n_rej = g["counters"]
    assert n_acc + n_rej == o["attempts"], tag


def _case(rows, cols, J, T, frac, seed, passes, period=0.0, occ=None):
    n = rows * cols
    if occ is None:
        occ = _occ(n, frac, seed)
    na = int(occ.sum())
    print(f"kmc case {rows}x{cols} J={J} T={T} atoms={na} period={period}", flush=True)
    with M.Run([dict(shape=(rows, cols), temperature=T, seed=seed)], mode=M.MODE_KMC, J1=J,
               sample_period=1, sample_time_period=period) as r:
        r.set_occupation(0, occ)
        r.reserve_samples(1 << 14)
        r.run(passes)
        g = _gpu_state(r, 0, na)
    o = O.kmc_run(rows, cols, J, T, passes * n, occ, O.mt64_state(seed, 0), time_period=period)
    _assert_equal(g, o, period, f"{rows}x{cols} J={J} T={T}")
    assert g["position"][:2] == (passes, 0)
    return g, o


@pytest.mark.parametrize("frac", [0.1, 0.5, 0.9])
@pytest.mark.parametrize("T", [300.0, 1000.0, 1e5])
def test_densities_and_temperatures(frac, T):
    _case(16, 12, 0.05, T, frac, seed=3, passes=8)


@pytest.mark.parametrize("J", [-0.04, 0.0])
def test_repulsive_and_free(J):
    _case(10, 14, J, 700.0, 0.4, seed=4, passes=10)


@pytest.mark.parametrize("period", [0.5, 3.0])
def test_time_grid(period):
    _case(12, 12, 0.05, 900.0, 0.3, seed=5, passes=6, period=period)


def test_small_and_extreme_lattices():
    _case(3, 3, 0.05, 800.0, 0.0, seed=6, passes=20, occ=np.array([1, 0, 0, 0, 1, 0, 0, 0, 0], np.int8))
    one = np.zeros(25, np.int8)
    one[7] = 1
    _case(5, 5, 0.1, 600.0, 0.0, seed=7, passes=30, occ=one)  # a random walker
    full = np.ones(30, np.int8)
    full[11] = 0
    _case(5, 6, 0.08, 1200.0, 0.0, seed=8, passes=10, occ=full)  # one vacancy


def test_largest_lattice():
    _case(75, 100, 0.05, 1000.0, 0.5, seed=9, passes=2)


def test_run_steps_chunks_and_passes():
    rows, cols, J, T = 9, 11, 0.06, 800.0
    n = rows * cols
    occ = _occ(n, 0.45, 10)
    na = int(occ.sum())
    with M.Run([dict(shape=(rows, cols), temperature=T, seed=10)], mode=M.MODE_KMC, J1=J,
               sample_period=1) as r:
        r.set_occupation(0, occ)
        total = 0
        for k in [1, 5, 60, n - 66, 2 * n + 7, 0, 13]:
            r.run_steps(k)
            total += k
            o = O.kmc_run(rows, cols, J, T, total, occ, O.mt64_state(10, 0))
            g = _gpu_state(r, 0, na)
            _assert_equal(g, o, 0.0, f"after {total} events")
            assert g["position"][:2] == (total // n, total % n)
        r.run(2)
        total += 2 * n
        o = O.kmc_run(rows, cols, J, T, total, occ, O.mt64_state(10, 0))
        _assert_equal(_gpu_state(r, 0, na), o, 0.0, "passes after events")


def test_time_grid_in_chunks():
    rows, cols, J, T, period = 8, 10, 0.05, 900.0, 0.75
    n = rows * cols
    occ = _occ(n, 0.35, 11)
    na = int(occ.sum())
    with M.Run([dict(shape=(rows, cols), temperature=T, seed=11)], mode=M.MODE_KMC, J1=J,
               sample_period=1, sample_time_period=period) as r:
        r.set_occupation(0, occ)
        r.reserve_samples(4096)
        for k in [3, 100, 1, 250]:
            r.run_steps(k)
        o = O.kmc_run(rows, cols, J, T, 354, occ, O.mt64_state(11, 0), time_period=period)
        _assert_equal(_gpu_state(r, 0, na), o, period, "time grid over four runs")


def test_batch_of_chains_equals_each_alone():
    shapes = [(6, 6), (12, 9), (20, 15), (3, 5), (30, 30)]
    Ts = [500.0, 900.0, 1500.0, 3000.0, 700.0]
    fracs = [0.2, 0.5, 0.7, 0.4, 0.3]
    chains = [dict(shape=s, temperature=T, seed=30 + k) for k, (s, T) in enumerate(zip(shapes, Ts))]
    occs = [_occ(s[0] * s[1], f, 40 + k) for k, (s, f) in enumerate(zip(shapes, fracs))]
    print("kmc batch of 5 chains", flush=True)
    with M.Run(chains, mode=M.MODE_KMC, J1=0.05, sample_period=1) as r:
        for k, o in enumerate(occs):
            r.set_occupation(k, o)
        r.run(5)
        for k, s in enumerate(shapes):
            o = O.kmc_run(s[0], s[1], 0.05, Ts[k], 5 * s[0] * s[1], occs[k],
                          O.mt64_state(30 + k, 0))
            _assert_equal(_gpu_state(r, k, int(occs[k].sum())), o, 0.0, f"chain {k}")


def test_frozen_chain_stops():
    # a 3x3 block of atoms: every hop costs 2 J or more, exp(-beta dE)
    # underflows at 10 K, so no attempt is ever accepted; a small spin limit
    # stops the chain
    rows = cols = 10
    occ = np.zeros(100, np.int8)
    for i in range(3, 6):
        for j in range(3, 6):
            occ[i + rows * j] = 1
    with M.Run([dict(shape=(rows, cols), temperature=10.0, seed=12)], mode=M.MODE_KMC, J1=0.5,
               sample_period=1) as r:
        r.set_occupation(0, occ)
        r.kmc_set_spin_limit(5000)
        with pytest.raises(M.McgError, match="spin limit"):
            r.run(1)
        assert np.array_equal(r.get_occupation(0), occ)
        assert r.counters(0)[1] == 0


def test_no_hop_possible_is_refused():
    with M.Run([dict(shape=(4, 4), temperature=1000.0)], mode=M.MODE_KMC, J1=0.05) as r:
        with pytest.raises(M.McgError, match="no hop"):
            r.run(1)
        r.set_occupation(0, np.ones(16, np.int8))
        with pytest.raises(M.McgError, match="no hop"):
            r.run(1)
    with pytest.raises(M.McgError, match="3x3"):
        M.Run([dict(shape=(2, 5), temperature=1000.0)], mode=M.MODE_KMC, J1=0.05)


def test_advance_selects_then_applies():
    # mcg_kmc_advance: the reference loop's order -- an event is selected (its
    # words drawn) before the samples, and applied after them
    rows, cols, J, T, seed = 7, 9, 0.06, 900.0, 13
    n = rows * cols
    occ = _occ(n, 0.4, seed)
    na = int(occ.sum())

    def ref(k, **kw):
        return O.kmc_run(rows, cols, J, T, k, occ, O.mt64_state(seed, 0), **kw)

    with M.Run([dict(shape=(rows, cols), temperature=T, seed=seed)], mode=M.MODE_KMC, J1=J,
               sample_period=1 << 40) as r:
        r.set_occupation(0, occ)
        r.kmc_advance(10, select_at_end=True)
        t, site, t_pending = r.nfold_state(0)
        o10, o11 = ref(10), ref(11)
        assert np.array_equal(r.get_occupation(0), o10["occ"])
        assert np.array_equal(r.rng_dump(0), o11["engine"])  # the 11th event's words are drawn
        assert t == o10["final_time"] and t_pending == o11["final_time"]
        moved_from = np.flatnonzero((o10["occ"] == 1) & (o11["occ"] == 0))
        assert list(moved_from) == [site]
        r.kmc_advance(5)  # the pending event, then four more
        o15 = ref(15)
        g = _gpu_state(r, 0, na)
        assert np.array_equal(g["occ"], o15["occ"]) and np.array_equal(g["atoms"], o15["atoms"])
        assert np.array_equal(g["engine"], o15["engine"]) and g["time"] == o15["final_time"]
        assert g["pending"] == -1 and g["position"][1] == 15
        # stop at a time: the event that reaches it is selected, not applied
        t_stop = o15["final_time"] + 3.0
        r.kmc_advance(1 << 40, stop_time=t_stop)
        grid = O.kmc_run(rows, cols, J, T, 5000, occ, O.mt64_state(seed, 0), time_period=t_stop)
        k = int(grid["events"][1])  # events applied before the one reaching t_stop
        ok, ok1 = ref(k), ref(k + 1)
        t, site, t_pending = r.nfold_state(0)
        assert np.array_equal(r.get_occupation(0), ok["occ"])
        assert np.array_equal(r.rng_dump(0), ok1["engine"])
        assert t == ok["final_time"] < t_stop <= t_pending == ok1["final_time"]
        assert site >= 0
        # an advance whose stop time the pending event reaches does nothing
        r.kmc_advance(100, stop_time=t_stop)
        assert np.array_equal(r.get_occupation(0), ok["occ"])
        # an ordinary run applies it first
        r.run_steps(7)
        o = ref(k + 7)
        assert np.array_equal(r.get_occupation(0), o["occ"])
        assert np.array_equal(r.rng_dump(0), o["engine"])
        assert r.nfold_state(0)[0] == o["final_time"]


@pytest.mark.parametrize("every,max_samples", [(1, 3), (3, 4)])
def test_hop_correlations_on_device(every, max_samples):
    """Hop correlations recorded by the kernel (mcg_kmc_collect_hops): every
    `every`-th jump of each atom, at most `max_samples` per atom, the atom's
    displacement, the event's step within its pass, its pass and its time --
    against the oracle replayed one event at a time (each prefix run shows
    which atom the event moved, its displacement, jumps and the time)."""
    rows, cols, J, T, seed, passes = 6, 5, 0.05, 900.0, 91, 5
    n = rows * cols
    occ = _occ(n, 0.4, 92)
    na = int(occ.sum())
    print(f"kmc hops every={every} max={max_samples} atoms={na}", flush=True)
    with M.Run([dict(shape=(rows, cols), temperature=T, seed=seed)], mode=M.MODE_KMC, J1=J,
               sample_period=1) as r:
        r.set_occupation(0, occ)
        r.kmc_collect_hops(every, max_samples)
        r.reserve_samples(64)
        r.run(passes)
        counts, rec, times = r.kmc_hops(0, na)
        g_atoms = r.atoms(0, na)
    want = [[] for _ in range(na)]
    prev = O.kmc_run(rows, cols, J, T, 0, occ, O.mt64_state(seed, 0))["atoms"]
    for e in range(1, passes * n + 1):
        o = O.kmc_run(rows, cols, J, T, e, occ, O.mt64_state(seed, 0))
        moved = np.nonzero(o["atoms"][:, 3] != prev[:, 3])[0]
        assert len(moved) == 1
        a = int(moved[0])
        _, dx, dy, jumps = o["atoms"][a]
        if jumps % every == 0 and len(want[a]) < max_samples:
            want[a].append((dx, dy, (e - 1) % n, (e - 1) // n, o["final_time"]))
        prev = o["atoms"]
    assert np.array_equal(g_atoms, prev)
    for a in range(na):
        assert counts[a] == len(want[a]), a
        for k, (dx, dy, st, ps, t) in enumerate(want[a]):
            assert tuple(rec[k, a]) == (dx, dy, st, ps), (a, k)
            assert times[k, a] == t, (a, k)
        assert np.all(rec[counts[a]:, a] == 0)
