"""CHECKERBOARD chains with an odd edge (ring colourings, checker_ring.hip).

A periodic lattice with an odd edge has no parity colouring; each axis is
coloured as a ring (x & 1, the last point of an odd ring colour 2), and a
site's class is the triple of its coordinates' colours.
Exact properties:
  - the device's integer observables equal the oracle's recount of the final
    lattice (2D and 3D, with and without J2, odd and mixed edges);
  - the LDS-resident and HBM kernels, and a chain alone or batched with
    parity-coloured chains, give identical trajectories.
Statistics (seeded, so deterministic):
  - exact enumeration of 3x5 and 5x5 (2D NN+NNN and NN) and 3x3x2 (3D);
  - agreement with the oracle's sequential random-site chain within 3 sigma.
"""
import math

import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O

pytestmark = pytest.mark.gpu


def run(chains, lattice, J2, passes, options=0, J1=0.1):
    r = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=lattice, J1=J1, J2=J2,
               options=options)
    r.run(passes)
    return r


@pytest.mark.parametrize("lattice,shape,J2", [
    (cm.LATTICE_SQUARE, (5, 7), -0.02),
    (cm.LATTICE_SQUARE, (9, 9), 0.0),
    (cm.LATTICE_SQUARE, (33, 16), -0.02),
    (cm.LATTICE_SQUARE, (3, 3), -0.02),
    (cm.LATTICE_SC, (3, 3, 5), -0.02),
    (cm.LATTICE_SC, (7, 6, 5), 0.0),
    (cm.LATTICE_SC, (4, 4, 9), -0.02),
])
@pytest.mark.parametrize("streaming", [0, cm.OPT_FORCE_STREAMING])
def test_counts_track_the_lattice(lattice, shape, J2, streaming):
    ndim = 3 if lattice == cm.LATTICE_SC else 2
    print("case: ring counts", shape, J2, streaming, flush=True)
    ch = [dict(shape=shape, temperature=2300.0, exchange_potential=0.04, seed=61)]
    r = run(ch, lattice, J2, 40, streaming)
    occ = r.get_occupation(0)
    S1, S2, M = O.lattice_observables(ndim, shape, occ)
    last = r.samples(0)[-1]
    assert (int(last["S1"]), int(last["M"])) == (S1, M)
    if J2:
        assert int(last["S2"]) == S2
    assert r.counters(0)[0] == 40 and r.counters(0)[1] > 0
    r.close()


@pytest.mark.parametrize("lattice,shapes", [
    (cm.LATTICE_SQUARE, [(33, 17), (16, 16), (5, 12), (64, 32), (7, 7)]),
    (cm.LATTICE_SC, [(9, 7, 5), (8, 8, 8), (3, 4, 4)]),
])
def test_lds_hbm_batched_and_alone_identical(lattice, shapes):
    """Ring-coloured chains beside parity-coloured ones: the LDS-resident launch,
    the HBM launches and each chain alone give one trajectory."""
    print("case: ring paths", lattice, shapes, flush=True)
    chains = [dict(shape=s, temperature=1900.0 + 150.0 * i, exchange_potential=0.03 * i,
                   seed=300 + i) for i, s in enumerate(shapes)]
    a = run(chains, lattice, -0.02, 23)
    b = run(chains, lattice, -0.02, 23, cm.OPT_FORCE_STREAMING)
    for i, ch in enumerate(chains):
        np.testing.assert_array_equal(a.get_occupation(i), b.get_occupation(i))
        np.testing.assert_array_equal(a.samples(i), b.samples(i))
        assert a.counters(i) == b.counters(i)
        alone = run([ch], lattice, -0.02, 23)
        np.testing.assert_array_equal(a.get_occupation(i), alone.get_occupation(0))
        np.testing.assert_array_equal(a.samples(i), alone.samples(0))
        alone.close()
    a.close()
    b.close()


def chain_means(r, n_chains, J1, J2, mu, N, drop):
    es, xs = [], []
    for c in range(n_chains):
        rec = r.samples(c)[drop:]
        Ef = -J1 * rec["S1"].astype(float) - J2 * rec["S2"].astype(float)
        Nx = (N + rec["M"]) / 2.0
        es.append((Ef - mu * Nx).mean())
        xs.append(Nx.mean())
    return np.array(es), np.array(xs)


@pytest.mark.parametrize("ndim,L,J2,T,mu", [
    (2, (3, 5), -0.02, 3000.0, 0.1),
    (2, (5, 5), -0.02, 3200.0, -0.05),
    (2, (5, 5), 0.0, 2800.0, 0.02),
    (3, (3, 3, 2), -0.02, 4000.0, 0.0),
])
def test_exact_enumeration(ndim, L, J2, T, mu):
    """<E> and <N x> of many independent seeded chains vs the exact SGC averages
    of every state of the lattice (4 standard errors)."""
    J1 = 0.1
    N = int(np.prod(L))
    ex = O.lattice_enumerate(ndim, L, J1, J2, T, mu)
    n_chains, passes, drop = 512, 4000, 500
    print("case: ring enumeration", ndim, L, J2, T, mu, flush=True)
    chains = [dict(shape=L, temperature=T, exchange_potential=mu, seed=5000 + i)
              for i in range(n_chains)]
    r = run(chains, cm.LATTICE_SC if ndim == 3 else cm.LATTICE_SQUARE, J2, passes)
    es, xs = chain_means(r, n_chains, J1, J2, mu, N, drop)
    r.close()
    for got, exact in ((es, ex["E"]), (xs, ex["Nx"])):
        se = got.std(ddof=1) / math.sqrt(n_chains)
        assert abs(got.mean() - exact) < 4 * se + 1e-12, (got.mean(), exact, se)


@pytest.mark.parametrize("ndim,L,J2,T,mu", [
    (2, (25, 25), -0.02, 2400.0, 0.05),
    (2, (31, 18), 0.0, 3300.0, -0.1),
    (3, (9, 9, 9), -0.02, 3500.0, -0.1),
])
def test_three_sigma_vs_oracle_sequential(ndim, L, J2, T, mu):
    """<e> and <x> of the ring-coloured chain agree with the oracle's sequential
    random-site chain within 3 sigma (BasicStatistics precision of each)."""
    J1 = 0.1
    N = int(np.prod(L))
    passes, drop = 6000, 500
    print("case: ring 3sigma", ndim, L, J2, T, mu, flush=True)
    _, S1, S2, M = O.lattice_run(ndim, L, J1, J2, T, mu, 11, passes, np.ones(N, dtype=np.int8))
    orec = np.zeros(passes, dtype=cm.SAMPLE_DTYPE)
    orec["S1"], orec["S2"], orec["M"] = S1, S2, M
    r = run([dict(shape=L, temperature=T, exchange_potential=mu, seed=12)],
            cm.LATTICE_SC if ndim == 3 else cm.LATTICE_SQUARE, J2, passes)
    grec = r.samples(0)
    r.close()
    for rec_o, rec_g in ((orec, grec),):
        for idx in (0, 1):
            def series(rec):
                Ef = -J1 * rec["S1"].astype(float) - J2 * rec["S2"].astype(float)
                Nx = (N + rec["M"]) / 2.0
                return ((Ef - mu * Nx) / N, Nx / N)[idx]
            ma, pa = O.basic_statistics(series(rec_o[drop:]))
            mb, pb = O.basic_statistics(series(rec_g[drop:]))
            sigma = math.hypot(pa, pb) / 1.959049
            assert abs(ma - mb) < 3 * sigma + 1e-12, (idx, ma, mb, sigma)
