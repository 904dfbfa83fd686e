"""The statistical-mode configurations at their own sizes on the GPU.

Parallel checkerboard (2D 4096^2 and 3D 256^3, J1 = 0.1, J2 = -0.02 eV, the
temperatures T_k = 1000 + 200 k K of the 16-temperature family):
  - two chains of each family for 3 passes: the 2D band kernels equal the
    per-colour kernels and the 4-site byte path; the 3D bit-plane slab kernels,
    the XCD-ordered line-class kernel, the LDS-tiled line-class kernel and the
    byte path are equal; the final integer observables equal the oracle's
    recount of the final lattice;
  - at a mid size (256^2, 32^3) and two of the family's temperatures, <e> and
    <x> agree with the oracle's sequential random-site chain within 3 sigma
    (BasicStatistics precision of each, combined in quadrature).
FCC cluster expansion batch (256 chains of 16^3..32^3 conventional cells in one
autotuned launch): every chain's counts equal the oracle's recount, and three
chains (16^3, 24^3, 32^3) equal the same chains run alone.
Served scan (the Poisson trace of 50/50 2D Ising / FCC requests, edges in
[16, 512] and [8, 48]): the first 200 requests with arrivals compressed 20x;
a sample including FCC edges >= 8, the FCC class above 32 cells (wide bit
lines) and the Ising HBM class, gives each request's result alone and in the
mix bit for bit.
"""
import math

import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O
from casmcode_monte_amd import fcc, serving

pytestmark = pytest.mark.gpu

J1, J2 = 0.1, -0.02


def family(shape, ks):
    return [dict(shape=shape, temperature=1000.0 + 200.0 * k, exchange_potential=0.0,
                 seed=0xC3000 + k) for k in ks]


def checker(chains, lattice, passes, options=0, **sched):
    r = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=lattice, J1=J1, J2=J2,
               options=options)
    if sched:
        r.set_schedule(**sched)
    r.run(passes)
    return r


def same_runs(a, b, n):
    for c in range(n):
        np.testing.assert_array_equal(a.get_occupation(c), b.get_occupation(c))
        np.testing.assert_array_equal(a.samples(c), b.samples(c))
        assert a.counters(c) == b.counters(c)


def test_c3_square_4096_paths_equal():
    chains = family((4096, 4096), (0, 15))
    print("case: c3 4096^2 band path", flush=True)
    band = checker(chains, cm.LATTICE_SQUARE, 3, cm.OPT_FORCE_STREAMING, band_lines=-1)
    assert band.schedule()["band_lines"] != 0
    print("case: c3 4096^2 per-colour path", flush=True)
    colour = checker(chains, cm.LATTICE_SQUARE, 3, cm.OPT_FORCE_STREAMING, band_lines=0)
    same_runs(band, colour, 2)
    colour.close()
    print("case: c3 4096^2 byte path", flush=True)
    quad = checker(chains, cm.LATTICE_SQUARE, 3,
                   cm.OPT_FORCE_STREAMING | cm.OPT_CHECKER_QUAD_PATH)
    same_runs(band, quad, 2)
    quad.close()
    for c in range(2):
        occ = band.get_occupation(c)
        S1, S2, M = O.lattice_observables(2, (4096, 4096), occ)
        last = band.samples(c)[-1]
        assert (int(last["S1"]), int(last["S2"]), int(last["M"])) == (S1, S2, M)
        assert band.counters(c)[0] == 3 and abs(M) < 4096 * 4096  # the lattice moved
    band.close()


def test_c3_cubic_256_paths_equal():
    chains = family((256, 256, 256), (0, 15))
    print("case: c3 256^3 slab path", flush=True)
    slab = checker(chains, cm.LATTICE_SC, 3, cm.OPT_FORCE_STREAMING, band_lines=-1)
    assert slab.schedule()["band_lines"] != 0
    print("case: c3 256^3 line-class path", flush=True)
    line = checker(chains, cm.LATTICE_SC, 3, cm.OPT_FORCE_STREAMING, band_lines=0, line_tiles=0)
    same_runs(slab, line, 2)
    slab.close()
    print("case: c3 256^3 tiled line-class path", flush=True)
    tiled = checker(chains, cm.LATTICE_SC, 3, cm.OPT_FORCE_STREAMING, band_lines=0, line_tiles=1)
    same_runs(line, tiled, 2)
    tiled.close()
    print("case: c3 256^3 byte path", flush=True)
    quad = checker(chains, cm.LATTICE_SC, 3, cm.OPT_FORCE_STREAMING | cm.OPT_CHECKER_QUAD_PATH)
    same_runs(line, quad, 2)
    quad.close()
    for c in range(2):
        occ = line.get_occupation(c)
        S1, S2, M = O.lattice_observables(3, (256, 256, 256), occ)
        last = line.samples(c)[-1]
        assert (int(last["S1"]), int(last["S2"]), int(last["M"])) == (S1, S2, M)
    line.close()


def energy_per_site(rec, mu, N):
    Ef = -J1 * rec["S1"].astype(float) - J2 * rec["S2"].astype(float)
    Nx = (N + rec["M"]) / 2.0
    return (Ef - mu * Nx) / N, Nx / N


@pytest.mark.parametrize("ndim,L,k", [
    (2, (256, 256), 2),     # 1400 K, ordered
    (2, (256, 256), 14),    # 3800 K, disordered
    (3, (32, 32, 32), 2),   # 1400 K
    (3, (32, 32, 32), 5),   # 2000 K
])
def test_c3_three_sigma_vs_oracle_mid_size(ndim, L, k):
    N = int(np.prod(L))
    T = 1000.0 + 200.0 * k
    passes, drop = 2000, 300
    print("case: c3 3sigma", ndim, L, T, flush=True)
    _, S1, S2, M = O.lattice_run(ndim, L, J1, J2, T, 0.0, 0xC3000 + k, passes,
                                 np.ones(N, dtype=np.int8))
    orec = np.zeros(passes, dtype=cm.SAMPLE_DTYPE)
    orec["S1"], orec["S2"], orec["M"] = S1, S2, M
    r = checker(family(L, (k,)), cm.LATTICE_SC if ndim == 3 else cm.LATTICE_SQUARE, passes)
    grec = r.samples(0)
    r.close()
    for idx in (0, 1):
        a = energy_per_site(orec[drop:], 0.0, N)[idx]
        b = energy_per_site(grec[drop:], 0.0, N)[idx]
        ma, pa = O.basic_statistics(a)
        mb, pb = O.basic_statistics(b)
        sigma = math.hypot(pa, pb) / 1.959049
        assert abs(ma - mb) < 3 * sigma + 1e-12, (idx, ma, mb, sigma)


def test_c4_fcc_batch_counts_and_alone():
    conds = fcc.batch_conditions()
    assert len(conds) == 256
    print("case: c4 fcc 256-chain batch", flush=True)
    r = cm.Run(conds, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_FCC, eci=fcc.DEFAULT_ECI)
    r.autotune(2)
    r.run(3)
    for c, ch in enumerate(conds):
        last = r.samples(c)[-1]
        got = dict(N1=int(last["M"]), P1=int(last["S1"]), P2=int(last["S2"]),
                   P3=int(last["S3"]), T=int(last["T3"]))
        assert got == O.fcc_counts(ch["shape"], r.get_occupation(c)), c
        assert r.counters(c)[0] == 3
    for c in (0, 8, 16):  # L = 16, 24, 32
        print("case: c4 fcc alone chain", c, flush=True)
        a = cm.Run([conds[c]], mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_FCC,
                   eci=fcc.DEFAULT_ECI)
        a.run(3)
        np.testing.assert_array_equal(a.get_occupation(0), r.get_occupation(c))
        np.testing.assert_array_equal(a.samples(0), r.samples(c))
        assert a.counters(0) == r.counters(c)
        a.close()
    r.close()


def result_key(res):
    return (res.passes, res.n_samples, res.converged, res.n_equilibration, res.mean_x,
            res.precision_x, res.lattice_crc)


def test_c5_served_trace_mix_equals_alone():
    reqs = serving.poisson_trace(200, rate=50.0, fcc_fraction=0.5, max_passes=3000)
    reqs = [serving.Request(**{**q.__dict__, "arrival": q.arrival / 20.0}) for q in reqs]
    print("case: c5 served mix of", len(reqs), "requests", flush=True)
    srv = serving.Server(fcc_classes=serving.DEFAULT_FCC_CLASSES)
    mixed, st = srv.serve(reqs)
    srv.close()
    assert len(mixed) == len(reqs)
    assert st["p50_latency_s"] <= st["p99_latency_s"]
    fcc_reqs = [q for q in reqs if q.model == "fcc"]
    ising = [q for q in reqs if q.model == "ising"]
    assert fcc_reqs and min(q.edge for q in fcc_reqs) >= 8
    # the sample: the largest FCC requests (the class above 32 cells), FCC
    # requests near the low end, the largest Ising requests (the 512 HBM class)
    # and a few small Ising requests
    by_edge = lambda qs: sorted(qs, key=lambda q: (q.edge, q.rid))
    sample = by_edge(fcc_reqs)[-3:] + by_edge(fcc_reqs)[:2] + by_edge(ising)[-3:] + \
        by_edge(ising)[:2]
    assert max(q.edge for q in sample if q.model == "fcc") > 32
    assert max(q.edge for q in sample if q.model == "ising") > 362
    for req in sample:
        print("case: c5 alone", req.rid, req.model, req.edge, flush=True)
        alone_srv = serving.Server(fcc_classes=serving.DEFAULT_FCC_CLASSES)
        alone, _ = alone_srv.serve([serving.Request(**{**req.__dict__, "arrival": 0.0})])
        alone_srv.close()
        assert result_key(alone[req.rid]) == result_key(mixed[req.rid]), req
