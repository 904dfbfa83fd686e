"""CHECKERBOARD mode on the GPU.

Bit-identity properties (the Philox counter depends only on chain seed, pass,
colour and site):
  - the LDS-resident kernel and the HBM work-list kernels give the same
    trajectory;
  - a chain run alone and the same chain run in a batch of uneven chains give
    the same trajectory.
Statistics (seeded, so deterministic):
  - exact enumeration of tiny lattices (2D NN+NNN, 3D NN+NNN);
  - Onsager's exact internal energy of the 2D nearest-neighbour model;
  - agreement within 3 sigma with the oracle's sequential random-site runs.
"""
import math

import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O

pytestmark = pytest.mark.gpu
KB = cm.KB


def energy_terms(rec, J1, J2, mu, N):
    Ef = -J1 * rec["S1"].astype(float) - J2 * rec["S2"].astype(float)
    Nx = (N + rec["M"]) / 2.0
    return Ef - mu * Nx, Nx


def run(chains, lattice, J1, J2, passes, options=0, sp=1):
    r = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=lattice, J1=J1, J2=J2,
               sample_period=sp, options=options)
    r.run(passes)
    return r


@pytest.mark.parametrize("lattice,shape,J2", [
    (cm.LATTICE_SQUARE, (16, 16), -0.02),
    (cm.LATTICE_SQUARE, (18, 10), 0.0),
    (cm.LATTICE_SC, (8, 6, 4), -0.02),
    (cm.LATTICE_SC, (6, 6, 6), 0.0),
    (cm.LATTICE_SC, (64, 6, 4), -0.02),
    (cm.LATTICE_SC, (32, 10, 6), 0.0),
])
def test_lds_and_streaming_paths_identical(lattice, shape, J2):
    print("case: lds vs streaming", lattice, shape, J2, flush=True)
    ch = [dict(shape=shape, temperature=2200.0, exchange_potential=0.05, seed=77)]
    a = run(ch, lattice, 0.1, J2, 40)
    b = run(ch, lattice, 0.1, J2, 40, options=cm.OPT_FORCE_STREAMING)
    np.testing.assert_array_equal(a.get_occupation(0), b.get_occupation(0))
    np.testing.assert_array_equal(a.samples(0), b.samples(0))
    assert a.counters(0) == b.counters(0)
    # integer observables track the lattice exactly
    S1, S2, M = O.lattice_observables(3 if lattice == cm.LATTICE_SC else 2, shape,
                                      a.get_occupation(0))
    last = a.samples(0)[-1]
    assert (last["S1"], last["M"]) == (S1, M)
    if J2 != 0.0:
        assert last["S2"] == S2


@pytest.mark.parametrize("streaming", [0, cm.OPT_FORCE_STREAMING])
def test_batched_equals_alone(streaming):
    chains = [dict(shape=(8, 8), temperature=1500.0, exchange_potential=0.2, seed=5),
              dict(shape=(32, 16), temperature=3000.0, exchange_potential=-0.1, seed=6),
              dict(shape=(20, 12), temperature=2600.0, exchange_potential=0.0, seed=7)]
    print("case: batched vs alone streaming", streaming, flush=True)
    batch = run(chains, cm.LATTICE_SQUARE, 0.1, -0.02, 30, options=streaming)
    for i, ch in enumerate(chains):
        alone = run([ch], cm.LATTICE_SQUARE, 0.1, -0.02, 30, options=streaming)
        np.testing.assert_array_equal(batch.get_occupation(i), alone.get_occupation(0))
        np.testing.assert_array_equal(batch.samples(i), alone.samples(0))


def chain_means(r, n_chains, J1, J2, mu, N, drop):
    es, xs = [], []
    for c in range(n_chains):
        rec = r.samples(c)[drop:]
        E, Nx = energy_terms(rec, J1, J2, mu, N)
        es.append(E.mean())
        xs.append(Nx.mean())
    return np.array(es), np.array(xs)


@pytest.mark.parametrize("ndim,L,J2,T,mu", [
    (2, (4, 4), -0.02, 3000.0, 0.1),
    (2, (4, 4), 0.0, 3500.0, -0.05),
    (2, (4, 6), 0.0, 2800.0, 0.02),
    (3, (2, 2, 4), -0.02, 4000.0, 0.0),
])
def test_exact_enumeration(ndim, L, J2, T, mu):
    """Many independent seeded chains per condition (temperatures above the
    tunnelling regime of these tiny cells, so 500 dropped passes equilibrate)."""
    J1 = 0.1
    N = int(np.prod(L))
    ex = O.lattice_enumerate(ndim, L, J1, J2, T, mu)
    n_chains, passes, drop = 512, 4000, 500
    print("case: enumeration", ndim, L, J2, T, mu, flush=True)
    chains = [dict(shape=L, temperature=T, exchange_potential=mu, seed=1000 + i)
              for i in range(n_chains)]
    r = run(chains, cm.LATTICE_SC if ndim == 3 else cm.LATTICE_SQUARE, J1, J2, passes)
    es, xs = chain_means(r, n_chains, J1, J2, mu, N, drop)
    for got, exact in ((es, ex["E"]), (xs, ex["Nx"])):
        se = got.std(ddof=1) / math.sqrt(n_chains)
        assert abs(got.mean() - exact) < 4 * se + 1e-12, (got.mean(), exact, se)


def onsager_energy_per_site(J, T):
    """Onsager (1944): u = -J coth(2K) [1 + (2/pi)(2 tanh^2(2K) - 1) K1(k^2)]."""
    from scipy.special import ellipk
    K = J / (KB * T)
    k = 2.0 * math.sinh(2 * K) / math.cosh(2 * K) ** 2
    return -J / math.tanh(2 * K) * (1 + 2 / math.pi * (2 * math.tanh(2 * K) ** 2 - 1)
                                    * ellipk(k * k))


@pytest.mark.parametrize("T", [1500.0, 4000.0])
def test_onsager_internal_energy(T):
    J1, L = 0.1, 128
    n_chains, passes, drop = 16, 2000, 300
    print("case: onsager T", T, flush=True)
    chains = [dict(shape=(L, L), temperature=T, exchange_potential=0.0, seed=50 + i)
              for i in range(n_chains)]
    r = run(chains, cm.LATTICE_SQUARE, J1, 0.0, passes)
    es, _ = chain_means(r, n_chains, J1, 0.0, 0.0, L * L, drop)
    u = es / (L * L)
    exact = onsager_energy_per_site(J1, T)
    se = u.std(ddof=1) / math.sqrt(n_chains)
    assert abs(u.mean() - exact) < 4 * se + 2e-5, (u.mean(), exact, se)


@pytest.mark.parametrize("ndim,L,J2,T,mu", [
    (2, (24, 24), -0.02, 2400.0, 0.05),
    (3, (8, 8, 8), -0.02, 3500.0, -0.1),
])
def test_three_sigma_vs_oracle_sequential(ndim, L, J2, T, mu):
    """<e> and <x> of the checkerboard chain agree with the oracle's sequential
    random-site chain within 3 sigma (sigma = BasicStatistics precision / 1.959049
    for each, combined in quadrature)."""
    J1 = 0.1
    N = int(np.prod(L))
    passes, drop = 6000, 500
    print("case: 3sigma", ndim, L, J2, T, mu, flush=True)
    occ0 = np.ones(N, dtype=np.int8)
    _, S1, S2, M = O.lattice_run(ndim, L, J1, J2, T, mu, 11, passes, occ0)
    orec = np.zeros(passes, dtype=cm.SAMPLE_DTYPE)
    orec["S1"], orec["S2"], orec["M"] = S1, S2, M
    r = run([dict(shape=L, temperature=T, exchange_potential=mu, seed=12)],
            cm.LATTICE_SC if ndim == 3 else cm.LATTICE_SQUARE, J1, J2, passes)
    grec = r.samples(0)
    for idx in (0, 1):
        a = energy_terms(orec[drop:], J1, J2, mu, N)[idx] / N
        b = energy_terms(grec[drop:], J1, J2, mu, N)[idx] / N
        ma, pa = O.basic_statistics(a)
        mb, pb = O.basic_statistics(b)
        sigma = math.hypot(pa, pb) / 1.959049
        assert abs(ma - mb) < 3 * sigma + 1e-12, (idx, ma, mb, sigma)


@pytest.mark.parametrize("lattice,shape,J2,streaming", [
    (cm.LATTICE_SQUARE, (64, 32), -0.02, 0),
    (cm.LATTICE_SQUARE, (32, 30), 0.0, cm.OPT_FORCE_STREAMING),
    (cm.LATTICE_SC, (32, 8, 6), -0.02, 0),
    (cm.LATTICE_SC, (64, 4, 4), 0.0, cm.OPT_FORCE_STREAMING),
])
def test_vector_path_equals_quad_path(lattice, shape, J2, streaming):
    """16-site vector (SWAR) path vs the 4-site byte path: identical trajectories."""
    print("case: chunk vs quad", lattice, shape, J2, streaming, flush=True)
    ch = [dict(shape=shape, temperature=2600.0, exchange_potential=-0.03, seed=91)]
    a = run(ch, lattice, 0.1, J2, 25, options=streaming)
    b = run(ch, lattice, 0.1, J2, 25, options=streaming | cm.OPT_CHECKER_QUAD_PATH)
    np.testing.assert_array_equal(a.get_occupation(0), b.get_occupation(0))
    np.testing.assert_array_equal(a.samples(0), b.samples(0))
    S1, S2, M = O.lattice_observables(3 if lattice == cm.LATTICE_SC else 2, shape,
                                      a.get_occupation(0))
    last = a.samples(0)[-1]
    assert (last["S1"], last["S2"] if J2 else S2, last["M"]) == (S1, S2, M)


def test_heat_capacity_and_susceptibility_vs_enumeration():
    """C = N var(e)/(kB T^2) and chi_x = N var(x)/(kB T) of many 4x4 J1+J2 chains
    vs the exact SGC ensemble (enumeration of all 2^16 states)."""
    from casmcode_monte_amd import analysis
    J1, J2, T, mu, L = 0.1, -0.02, 3200.0, 0.05, (4, 4)
    N = 16
    ex = O.lattice_enumerate(2, L, J1, J2, T, mu)
    C_exact = (ex["E2"] - ex["E"] ** 2) / (N * KB * T * T)
    chi_exact = (ex["Nx2"] - ex["Nx"] ** 2) / (N * KB * T)
    n_chains, passes, drop = 512, 4000, 500
    print("case: fluctuations vs enumeration", flush=True)
    chains = [dict(shape=L, temperature=T, exchange_potential=mu, seed=4000 + i)
              for i in range(n_chains)]
    r = run(chains, cm.LATTICE_SQUARE, J1, J2, passes)
    Cs, chis = [], []
    for c in range(n_chains):
        obs = r.observables(c)
        Cs.append(analysis.heat_capacity(obs["potential_energy"][drop:], N, T)[0])
        chis.append(analysis.susceptibility(obs["param_composition"][drop:], N, T)[0])
    for vals, exact in ((np.array(Cs), C_exact), (np.array(chis), chi_exact)):
        se = vals.std(ddof=1) / math.sqrt(n_chains)
        assert abs(vals.mean() - exact) < 4 * se + 0.01 * abs(exact), (vals.mean(), exact, se)


SCHED_CHAINS = [dict(shape=(64, 48), temperature=2400.0, exchange_potential=0.05, seed=31),
                dict(shape=(40, 36), temperature=1800.0, exchange_potential=-0.1, seed=32)]


@pytest.mark.parametrize("streaming,schedules", [
    (0, [dict(lds_block=256), dict(lds_block=512), dict(lds_block=1024)]),
    (cm.OPT_FORCE_STREAMING, [dict(units_per_thread=u, graph_passes=g)
                              for u in (1, 2, 4) for g in (1, 8, 5)]),
])
def test_schedules_give_identical_trajectories(streaming, schedules):
    """Launch schedules (LDS block size, units per thread, graph replay of
    several passes) change nothing but speed; 37 passes leave a remainder
    after the graph replays."""
    print("case: schedules streaming", streaming, flush=True)
    ref = None
    for s in schedules:
        r = cm.Run(SCHED_CHAINS, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_SQUARE,
                   J1=0.1, J2=-0.02, options=streaming)
        r.set_schedule(**s)
        r.run(37)
        r.run(11)
        got = [(r.get_occupation(c), r.samples(c), r.counters(c)) for c in range(2)]
        if ref is None:
            ref = got
            continue
        for (oa, sa, ca), (ob, sb, cb) in zip(ref, got):
            np.testing.assert_array_equal(oa, ob)
            np.testing.assert_array_equal(sa, sb)
            assert ca == cb


@pytest.mark.parametrize("streaming", [0, cm.OPT_FORCE_STREAMING])
def test_autotune_leaves_the_run_unchanged(streaming):
    print("case: autotune streaming", streaming, flush=True)
    a = cm.Run(SCHED_CHAINS, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_SQUARE,
               J1=0.1, J2=-0.02, options=streaming)
    b = cm.Run(SCHED_CHAINS, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_SQUARE,
               J1=0.1, J2=-0.02, options=streaming)
    a.run(10)
    b.run(10)
    chosen = a.autotune(24)
    assert chosen["trial_ms"] > 0.0
    assert a.schedule()["units_per_thread"] == chosen["units_per_thread"]
    a.run(30)
    b.run(30)
    for c in range(2):
        np.testing.assert_array_equal(a.get_occupation(c), b.get_occupation(c))
        np.testing.assert_array_equal(a.samples(c), b.samples(c))
        assert a.counters(c) == b.counters(c)
    # a second handle of the same size class hits the cache
    assert b.autotune(24) == chosen


@pytest.mark.parametrize("J2", [0.0, -0.02])
@pytest.mark.parametrize("band_lines", [2, 4, 8, 64, -1])
def test_band_path_equals_colour_kernels(J2, band_lines):
    """The 2D two-launch band kernels (bands of band_lines lines, including one
    band per chain and bands of 2 lines) give the per-colour kernels'
    trajectories bit for bit, and those equal the LDS-resident kernel's."""
    print("case: band path", J2, band_lines, flush=True)
    chains = [dict(shape=(64, 24), temperature=2300.0, exchange_potential=0.04, seed=71),
              dict(shape=(32, 2), temperature=1700.0, exchange_potential=-0.2, seed=72),
              dict(shape=(96, 10), temperature=3100.0, exchange_potential=0.0, seed=73),
              dict(shape=(32, 64), temperature=900.0, exchange_potential=0.1, seed=74)]
    ref = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_SQUARE, J1=0.1, J2=J2)
    ref.run(21)
    for G in (1, 8):
        for bl in (0, band_lines):
            r = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_SQUARE, J1=0.1,
                       J2=J2, options=cm.OPT_FORCE_STREAMING)
            r.set_schedule(band_lines=bl, graph_passes=G)
            r.run(21)
            for c in range(len(chains)):
                np.testing.assert_array_equal(ref.get_occupation(c), r.get_occupation(c))
                np.testing.assert_array_equal(ref.samples(c), r.samples(c))
                assert ref.counters(c) == r.counters(c)


@pytest.mark.parametrize("band_lines", [2, 4, 6, -1])
def test_slab_path_equals_other_paths(band_lines):
    """The 3D J1+J2 slab kernels (planes bit-packed in an LDS ring, slabs of
    band_lines planes, the last plane of each slab in the edge launch) give the
    LDS-resident kernel's and the line-class kernel's trajectories bit for bit,
    for a batch of uneven chains (rows of 1 to 3 words per half, 2 to 12 planes,
    slabs that hold a whole chain, planes of two rows)."""
    print("case: slab path", band_lines, flush=True)
    chains = [dict(shape=(64, 6, 4), temperature=2300.0, exchange_potential=0.04, seed=81),
              dict(shape=(128, 2, 2), temperature=1700.0, exchange_potential=-0.2, seed=82),
              dict(shape=(192, 10, 12), temperature=3100.0, exchange_potential=0.0, seed=83),
              dict(shape=(64, 4, 6), temperature=900.0, exchange_potential=0.1, seed=84)]
    ref = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_SC, J1=0.1, J2=-0.02)
    ref.run(17)
    line = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_SC, J1=0.1, J2=-0.02,
                  options=cm.OPT_FORCE_STREAMING)
    line.set_schedule(band_lines=0)
    line.run(17)
    for G in (1, 8):
        r = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_SC, J1=0.1, J2=-0.02,
                   options=cm.OPT_FORCE_STREAMING)
        r.set_schedule(band_lines=band_lines, graph_passes=G)
        r.run(17)
        for c in range(len(chains)):
            for other in (ref, line):
                np.testing.assert_array_equal(other.get_occupation(c), r.get_occupation(c))
                np.testing.assert_array_equal(other.samples(c), r.samples(c))
                assert other.counters(c) == r.counters(c)
            S1, S2, M = O.lattice_observables(3, chains[c]["shape"], r.get_occupation(c))
            last = r.samples(c)[-1]
            assert (int(last["S1"]), int(last["S2"]), int(last["M"])) == (S1, S2, M)
        r.close()


@pytest.mark.parametrize("lattice,shape", [
    (cm.LATTICE_SC, (64, 10, 6)),
    (cm.LATTICE_SC, (32, 14, 10)),
    (cm.LATTICE_SC, (96, 6, 18)),
    (cm.LATTICE_SC, (32, 2, 2)),
    (cm.LATTICE_SQUARE, (64, 30)),
    (cm.LATTICE_SQUARE, (128, 6)),
])
def test_line_class_tiles_equal_other_paths(lattice, shape):
    """The LDS-tiled line-class kernel (tiles sized by units_per_thread, with
    partial tiles at the lattice edges and wrap-around neighbour lines) gives
    the LDS-resident kernel's and the untiled line-class kernel's trajectories
    bit for bit."""
    print("case: line-class tiles", lattice, shape, flush=True)
    ch = [dict(shape=shape, temperature=2500.0, exchange_potential=0.03, seed=123),
          dict(shape=shape, temperature=1400.0, exchange_potential=-0.1, seed=124)]
    ref = run(ch, lattice, 0.1, -0.02, 19)
    quad = run(ch, lattice, 0.1, -0.02, 19,
               options=cm.OPT_FORCE_STREAMING | cm.OPT_CHECKER_QUAD_PATH)
    for u in (1, 2, 4):
        for G in (1, 8):
            r = cm.Run(ch, mode=cm.MODE_CHECKERBOARD, lattice=lattice, J1=0.1, J2=-0.02,
                       options=cm.OPT_FORCE_STREAMING)
            r.set_schedule(units_per_thread=u, graph_passes=G, line_tiles=1, band_lines=0)
            r.run(19)
            for c in range(2):
                for other in (ref, quad):
                    np.testing.assert_array_equal(other.get_occupation(c), r.get_occupation(c))
                    np.testing.assert_array_equal(other.samples(c), r.samples(c))
                    assert other.counters(c) == r.counters(c)


@pytest.mark.parametrize("ndim,L,J2,T,mu", [
    (2, (32, 32), -0.02, 3000.0, 0.05),
    (2, (24, 24), 0.0, 3600.0, -0.1),
    (3, (10, 10, 10), -0.02, 4200.0, 0.0),
])
def test_fluctuations_three_sigma_vs_oracle_sequential(ndim, L, J2, T, mu):
    """Heat capacity C = N var(e) / (kB T^2) and susceptibility chi_x = N var(x) /
    (kB T) of the checkerboard chain and of the oracle's sequential random-site
    chain (the reference's CPU path) agree within 3 sigma, sigma from the
    BasicStatistics precision of the squared deviations of each, combined in
    quadrature."""
    from casmcode_monte_amd import analysis
    J1 = 0.1
    N = int(np.prod(L))
    passes, drop = 8000, 500
    print("case: fluctuations 3sigma", ndim, L, J2, T, mu, flush=True)
    _, S1, S2, M = O.lattice_run(ndim, L, J1, J2, T, mu, 21, passes, np.ones(N, dtype=np.int8))
    orec = np.zeros(passes, dtype=cm.SAMPLE_DTYPE)
    orec["S1"], orec["S2"], orec["M"] = S1, S2, M
    r = run([dict(shape=L, temperature=T, exchange_potential=mu, seed=22)],
            cm.LATTICE_SC if ndim == 3 else cm.LATTICE_SQUARE, J1, J2, passes)
    grec = r.samples(0)
    r.close()

    def fluct(rec):
        E, Nx = energy_terms(rec[drop:], J1, J2, mu, N)
        return (analysis.heat_capacity_stats(E / N, N, T),
                analysis.susceptibility_stats(Nx / N, N, T))

    for (va, pa), (vb, pb) in zip(fluct(orec), fluct(grec)):
        sigma = math.hypot(pa, pb) / 1.959964
        assert va > 0 and vb > 0
        assert abs(va - vb) < 3 * sigma, (va, vb, sigma)
