"""Binary FCC cluster expansion on the GPU (CHECKERBOARD, 64 colour classes).

Exact properties:
  - the device counts (N1, P1, P2, P3, T) equal the oracle's recount of the
    final lattice (odd and even edges, so both colourings run);
  - the LDS-resident and the global-memory kernels, every block size, and a
    chain alone or in a batch give identical trajectories.
Statistics (seeded, so deterministic):
  - non-interacting sites: <x> = 1/(1 + exp((V1 - mu)/kT)) within 4 sigma;
  - <e> and <x> agree with the oracle's sequential random-site chain within
    3 sigma (BasicStatistics precision of each, combined in quadrature).
"""
import math

import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O
from casmcode_monte_amd import fcc

pytestmark = pytest.mark.gpu
ECI = fcc.DEFAULT_ECI


def run(chains, passes, options=0, eci=ECI, block=None, line_tiles=None):
    r = cm.Run(chains, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_FCC, eci=eci,
               options=options)
    if block or line_tiles is not None:
        r.set_schedule(lds_block=block, line_tiles=line_tiles)
    r.run(passes)
    return r


def counts_of(rec):
    return dict(N1=int(rec["M"]), P1=int(rec["S1"]), P2=int(rec["S2"]), P3=int(rec["S3"]),
                T=int(rec["T3"]))


@pytest.mark.parametrize("shape,T,mu", [
    ((3, 3, 3), 900.0, -0.05),
    ((4, 4, 4), 1500.0, 0.1),
    ((5, 3, 4), 700.0, -0.2),
    ((6, 6, 6), 2000.0, 0.3),
])
def test_counts_track_the_lattice(shape, T, mu):
    print("case: fcc counts", shape, T, mu, flush=True)
    r = run([dict(shape=shape, temperature=T, exchange_potential=mu, seed=5)], 40)
    occ = r.get_occupation(0)
    assert set(np.unique(occ)) <= {0, 1}
    assert counts_of(r.samples(0)[-1]) == O.fcc_counts(shape, occ)
    n_pass, n_acc, _ = r.counters(0)
    assert n_pass == 40 and n_acc > 0


@pytest.mark.parametrize("shape", [(4, 4, 4), (5, 3, 4)])
def test_lds_global_and_block_sizes_identical(shape):
    print("case: fcc lds/global/block", shape, flush=True)
    ch = [dict(shape=shape, temperature=1100.0, exchange_potential=-0.02, seed=17)]
    ref = run(ch, 25)
    for opts, block in ((cm.OPT_FORCE_STREAMING, None), (0, 256), (0, 512),
                        (cm.OPT_FORCE_STREAMING, 256)):
        got = run(ch, 25, options=opts, block=block)
        np.testing.assert_array_equal(ref.get_occupation(0), got.get_occupation(0))
        np.testing.assert_array_equal(ref.samples(0), got.samples(0))
        assert ref.counters(0) == got.counters(0)


@pytest.mark.parametrize("shapes", [
    [(40, 6, 8)],
    [(33, 4, 4), (47, 5, 7), (64, 4, 6)],
    [(48, 10, 10), (16, 6, 4), (5, 3, 4)],  # narrow chains in a launch of wide lines
])
def test_wide_bit_lines_equal_global_kernel(shapes):
    """Lines of 33 to 64 sites in the bit-line kernel (the line once per word,
    doubled in registers, 512-thread workgroups): the trajectories equal the
    global-memory byte kernel's and the counts the oracle's recount, for
    chains of both line lengths in one launch and two block sizes."""
    print("case: fcc wide bit lines", shapes, flush=True)
    chains = [dict(shape=sh, temperature=900.0 + 300.0 * i, exchange_potential=0.05 - 0.1 * i,
                   seed=30 + i) for i, sh in enumerate(shapes)]
    ref = run(chains, 12, options=cm.OPT_FORCE_STREAMING)
    for block in (None, 256):
        got = run(chains, 12, block=block)
        for c in range(len(chains)):
            np.testing.assert_array_equal(ref.get_occupation(c), got.get_occupation(c))
            np.testing.assert_array_equal(ref.samples(c), got.samples(c))
            assert ref.counters(c) == got.counters(c)
            assert counts_of(got.samples(c)[-1]) == O.fcc_counts(chains[c]["shape"],
                                                                  got.get_occupation(c))
        got.close()
    ref.close()


@pytest.mark.parametrize("shapes", [
    [(3, 3, 3)],
    [(5, 3, 4), (7, 9, 5), (4, 4, 4)],
    [(32, 4, 6), (17, 16, 18), (31, 5, 3), (16, 16, 16)],
])
def test_class_line_kernel_equals_unit_kernels(shapes):
    """The class-line kernel (one thread per class line, bit-sliced counts;
    schedule line_tiles = 0, the default for lines of at most 32 sites) gives
    the trajectories of the 4-site unit bit-line kernel (line_tiles = 1) and of
    the global-memory byte kernel, for even and odd edges (whose colourings
    put sites of one colour on lines of both parities), and counts equal to
    the oracle's recount."""
    print("case: fcc class lines", shapes, flush=True)
    chains = [dict(shape=sh, temperature=700.0 + 400.0 * i, exchange_potential=0.1 - 0.15 * i,
                   seed=60 + i) for i, sh in enumerate(shapes)]
    lines = run(chains, 15)
    assert lines.schedule()["line_tiles"] == 0
    for other in (run(chains, 15, line_tiles=1, block=512),
                  run(chains, 15, options=cm.OPT_FORCE_STREAMING)):
        for c in range(len(chains)):
            np.testing.assert_array_equal(lines.get_occupation(c), other.get_occupation(c))
            np.testing.assert_array_equal(lines.samples(c), other.samples(c))
            assert lines.counters(c) == other.counters(c)
        other.close()
    for c in range(len(chains)):
        assert counts_of(lines.samples(c)[-1]) == O.fcc_counts(chains[c]["shape"],
                                                                lines.get_occupation(c))
        assert lines.counters(c)[1] > 0
    lines.close()


def test_batched_equals_alone():
    chains = [dict(shape=(3, 3, 3), temperature=800.0, exchange_potential=0.05, seed=1),
              dict(shape=(6, 4, 5), temperature=1700.0, exchange_potential=-0.3, seed=2),
              dict(shape=(4, 4, 4), temperature=400.0, exchange_potential=0.0, seed=3)]
    print("case: fcc batched vs alone", flush=True)
    batch = run(chains, 30)
    for i, ch in enumerate(chains):
        alone = run([ch], 30)
        np.testing.assert_array_equal(batch.get_occupation(i), alone.get_occupation(0))
        np.testing.assert_array_equal(batch.samples(i), alone.samples(0))


def test_set_occupation_and_assign():
    print("case: fcc set occupation / assign", flush=True)
    shape = (4, 3, 5)
    N = 4 * 60
    r = cm.Run([dict(shape=shape, temperature=900.0, seed=8)], mode=cm.MODE_CHECKERBOARD,
               lattice=cm.LATTICE_FCC, eci=ECI)
    occ = (np.random.default_rng(1).random(N) < 0.5).astype(np.int8)
    r.set_occupation(0, occ)
    np.testing.assert_array_equal(r.get_occupation(0), occ)
    r.run(5)
    assert counts_of(r.samples(0)[-1]) == O.fcc_counts(shape, r.get_occupation(0))
    r.assign(0, dict(shape=(3, 3, 3), temperature=1200.0, exchange_potential=0.1, seed=9))
    assert r.n_samples(0) == 0 and not r.get_occupation(0).any()
    r.run(5)
    assert counts_of(r.samples(0)[-1]) == O.fcc_counts((3, 3, 3), r.get_occupation(0))


@pytest.mark.parametrize("T,mu", [(1000.0, -0.05), (600.0, -0.12)])
def test_noninteracting_closed_form(T, mu):
    print("case: fcc non-interacting", T, mu, flush=True)
    V1 = -0.1
    eci = (0.0, V1, 0.0, 0.0, 0.0, 0.0)
    n_chains, passes, drop = 32, 1500, 100
    chains = [dict(shape=(4, 4, 4), temperature=T, exchange_potential=mu, seed=300 + i)
              for i in range(n_chains)]
    r = run(chains, passes, eci=eci)
    xs = np.array([(r.samples(c)["M"][drop:] / 256.0).mean() for c in range(n_chains)])
    exact = 1.0 / (1.0 + math.exp((V1 - mu) / (cm.KB * T)))
    se = xs.std(ddof=1) / math.sqrt(n_chains)
    assert abs(xs.mean() - exact) < 4 * se + 1e-12, (xs.mean(), exact, se)


@pytest.mark.parametrize("shape,T,mu", [
    ((3, 3, 3), 1500.0, -0.1),
    ((4, 4, 4), 900.0, -0.08),
    ((4, 4, 4), 2000.0, 0.2),
])
def test_three_sigma_vs_oracle_sequential(shape, T, mu):
    print("case: fcc 3sigma", shape, T, mu, flush=True)
    N = 4 * shape[0] * shape[1] * shape[2]
    passes, drop = 6000, 500
    _, oc = O.fcc_run(shape, ECI, T, mu, 21, passes, np.zeros(N))
    orec = np.zeros(passes, dtype=cm.SAMPLE_DTYPE)
    orec["M"], orec["S1"], orec["S2"], orec["S3"], orec["T3"] = oc.T
    r = run([dict(shape=shape, temperature=T, exchange_potential=mu, seed=22)], passes)
    og = r.observables(0, orec)
    gg = r.observables(0)
    for key in ("potential_energy", "param_composition"):
        ma, pa = O.basic_statistics(og[key][drop:])
        mb, pb = O.basic_statistics(gg[key][drop:])
        sigma = math.hypot(pa, pb) / 1.959049
        assert abs(ma - mb) < 3 * sigma + 1e-12, (key, ma, mb, sigma)


def test_device_statistics_match_host():
    print("case: fcc device stats", flush=True)
    r = run([dict(shape=(4, 4, 4), temperature=1200.0, exchange_potential=0.0, seed=4)], 800)
    for obs, key in ((cm.OBS_FORMATION_ENERGY, "formation_energy"),
                     (cm.OBS_PARAM_COMPOSITION, "param_composition"),
                     (cm.OBS_POTENTIAL_ENERGY, "potential_energy")):
        st = r.stats(0, obs, first=100)
        m, prec = O.basic_statistics(r.observables(0)[key][100:])
        assert abs(st["mean"] - m) <= 1e-12 * max(1.0, abs(m))
        assert abs(st["calculated_precision"] - prec) <= 1e-9 * max(1e-12, abs(prec)) + 1e-15
