"""EXACT mode on the GPU vs the oracle: bit-exact trajectories.

The oracle (oracle/refcpu.hh) runs the reference loop of
basic_occupation_metropolis.hh:381-422 with std::mt19937_64, the libstdc++
distributions and std::exp; the GPU kernel must reproduce every sampled
integer observable, every sampled FP64 value (bitwise), the acceptance counts,
the final occupation and the final engine state words.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import casmcode_monte_amd as cm
import oracle as O

pytestmark = pytest.mark.gpu


def run_gpu(rows, cols, T, mu, seed, passes, use_nlist=True, sample_period=1, J=0.1,
            init=None, chunks=None, options=0):
    r = cm.Run([dict(shape=(rows, cols), temperature=T, exchange_potential=mu, seed=seed)],
               mode=cm.MODE_EXACT, J1=J, use_nlist=use_nlist, sample_period=sample_period,
               options=options)
    if init is not None:
        r.set_occupation(0, init)
    for k in (chunks or [passes]):
        r.run(k)
    return r


def compare(r, o, c=0):
    rec = r.samples(c)
    assert len(rec) == o["n_samples"]
    np.testing.assert_array_equal(rec["S1"], o["S"])
    np.testing.assert_array_equal(rec["M"], o["M"])
    obs = r.observables(c, rec)
    for k in ("formation_energy", "param_composition", "potential_energy"):
        np.testing.assert_array_equal(obs[k].view(np.uint64), o[k].view(np.uint64), err_msg=k)
    n_pass, n_acc, n_rej = r.counters(c)
    assert (n_pass, n_acc, n_rej) == (o["n_pass"], o["n_accept"], o["n_reject"])
    np.testing.assert_array_equal(r.get_occupation(c).astype(np.int32), o["final_occ"])
    np.testing.assert_array_equal(r.rng_dump(c), o["engine"])


@pytest.mark.parametrize("use_nlist", [True, False])
def test_c1_plain_reference_run(use_nlist):
    """25x25, J=0.1, mu=0, T=2000 K, all +1, seed 42, 11000 passes sampled every pass."""
    print("case: C1 25x25 seed 42 use_nlist", use_nlist, flush=True)
    o = O.sgc_run(25, 25, J=0.1, use_nlist=use_nlist, temperature=2000.0, mu=0.0, seed=42,
                  max_count=11000)
    r = run_gpu(25, 25, 2000.0, 0.0, 42, 11000, use_nlist=use_nlist)
    compare(r, o)


@pytest.mark.parametrize("shape,T,mu,seed,sp", [
    ((7, 13), 600.0, 0.15, 7, 1),
    ((16, 16), 5000.0, -0.3, 123, 3),
    ((1, 40), 2500.0, 0.05, 2, 1),
    ((32, 8), 1200.0, -0.4, 99, 2),
])
@pytest.mark.parametrize("use_nlist", [True, False])
def test_exact_variants(shape, T, mu, seed, sp, use_nlist):
    rows, cols = shape
    passes = 300
    print("case:", shape, T, mu, seed, sp, use_nlist, flush=True)
    rng = np.random.default_rng(seed)
    init = rng.choice([-1, 1], size=rows * cols).astype(np.int32)
    o = O.sgc_run(rows, cols, temperature=T, mu=mu, seed=seed, max_count=passes,
                  sample_period=sp, init_occ=init, use_nlist=use_nlist)
    r = run_gpu(rows, cols, T, mu, seed, passes, use_nlist=use_nlist, sample_period=sp,
                init=init.astype(np.int8), chunks=[1, 120, 179])
    compare(r, o)
    # the running potential energy (occupation_metropolis.hh:106,141) stays within
    # accumulated rounding of the recomputed per-unitcell potential (not on a
    # 1-wide lattice: there the reference's dE counts the site as its own
    # neighbour, model.hh:312-314, while its per_supercell does not change)
    if rows > 1 and cols > 1:
        rec = r.samples(0)
        obs = r.observables(0, rec)
        np.testing.assert_allclose(rec["e_run"], obs["potential_energy"], rtol=0, atol=1e-9)


def test_rng_dump_load_roundtrip():
    """Dump the engine after a run, load it into a fresh handle: same continuation."""
    a = run_gpu(12, 12, 1800.0, 0.1, 5, 40)
    occ = a.get_occupation(0)
    words = a.rng_dump(0)
    a.run(25)
    b = run_gpu(12, 12, 1800.0, 0.1, 5, 0, init=occ)
    b.rng_load(0, words)
    b.run(25)
    np.testing.assert_array_equal(a.get_occupation(0), b.get_occupation(0))
    np.testing.assert_array_equal(a.rng_dump(0), b.rng_dump(0))
    np.testing.assert_array_equal(a.samples(0)[-25:]["S1"], b.samples(0)["S1"])


def c2_conditions():
    """Batched exact scan: T = 500 + a*2500/7, mu = -0.4 + b*0.8/7, seed 1000+i."""
    out = []
    for a in range(8):
        for b in range(8):
            out.append(dict(shape=(128, 128), temperature=500.0 + a * 2500.0 / 7.0,
                            exchange_potential=-0.4 + b * 0.8 / 7.0, seed=1000 + 8 * a + b))
    return out


def test_c2_batched_64_chains_bit_exact():
    """All 64 chains of the 128x128 scan in one launch; 60 passes each vs the oracle."""
    passes = 60
    conds = c2_conditions()
    print("case: C2 64 chains x 128^2, passes", passes, flush=True)
    r = cm.Run(conds, mode=cm.MODE_EXACT, J1=0.1, use_nlist=True)
    r.run(passes)

    def orc(c):
        return O.sgc_run(128, 128, temperature=c["temperature"], mu=c["exchange_potential"],
                         seed=c["seed"], max_count=passes)

    with cf.ThreadPoolExecutor(8) as ex:
        outs = list(ex.map(orc, conds))
    for i, o in enumerate(outs):
        compare(r, o, c=i)


@pytest.mark.parametrize("shape,T,mu", [((128, 128), 2200.0, 0.1), ((9, 11), 3000.0, -0.2),
                                         ((256, 64), 4000.0, 0.0), ((2, 50), 1500.0, 0.3)])
def test_speculative_walk_equals_scalar_walk(shape, T, mu):
    """The block-speculative walk and the one-attempt scalar walk consume the
    engine identically (so the rare Lemire-rejection path, which is the scalar
    walk's code, is exercised against the same oracle trajectories)."""
    rows, cols = shape
    print("case: speculative vs scalar", shape, T, mu, flush=True)
    a = run_gpu(rows, cols, T, mu, 31, 0)
    b = run_gpu(rows, cols, T, mu, 31, 0, options=cm.OPT_EXACT_SCALAR_WALK)
    for k in (3, 1, 40):
        a.run(k)
        b.run(k)
        np.testing.assert_array_equal(a.get_occupation(0), b.get_occupation(0))
        np.testing.assert_array_equal(a.rng_dump(0), b.rng_dump(0))
        np.testing.assert_array_equal(a.samples(0), b.samples(0))
        assert a.counters(0) == b.counters(0)


def test_large_lattice_vs_oracle():
    """256x256 (window 64, many blocks per pass) bit-exact for 12 passes."""
    print("case: 256x256 vs oracle", flush=True)
    o = O.sgc_run(256, 256, temperature=2600.0, mu=0.02, seed=77, max_count=12)
    r = run_gpu(256, 256, 2600.0, 0.02, 77, 12)
    compare(r, o)


@pytest.mark.parametrize("shape,T,seed", [((4, 4), 3000.0, 17), ((2, 6), 800.0, 18),
                                          ((10, 10), 1500.0, 19)])
def test_many_short_launches(shape, T, seed):
    """Many one- and two-pass launches: each launch starts and ends the engine
    hand-off between the walking and the twisting wavefront, at every offset
    into a generation; the trajectory must equal one uninterrupted run."""
    rows, cols = shape
    passes = 90
    print("case: short launches", shape, T, seed, flush=True)
    o = O.sgc_run(rows, cols, J=0.1, use_nlist=True, temperature=T, mu=0.05, seed=seed,
                  max_count=passes)
    r = run_gpu(rows, cols, T, 0.05, seed, passes, chunks=[1, 2] * 30)
    compare(r, o)
    r.close()


@pytest.mark.parametrize("shape,use_nlist", [((360, 400), True), ((380, 380), False)])
def test_lattice_at_the_lds_limit(shape, use_nlist):
    """Lattices that just fit the block walk's LDS (144000 and 144400 lattice
    bytes plus the fold scratch, at most 147456 bytes, beside the walk's ~15 KB
    of engine generations): the dynamic-LDS attribute must admit them;
    bit-exact for 2 passes."""
    rows, cols = shape
    print("case: LDS limit", shape, use_nlist, flush=True)
    o = O.sgc_run(rows, cols, use_nlist=use_nlist, temperature=2400.0, mu=0.0, seed=5,
                  max_count=2)
    r = run_gpu(rows, cols, 2400.0, 0.0, 5, 2, use_nlist=use_nlist)
    compare(r, o)
    r.close()


@pytest.mark.parametrize("use_nlist", [True, False])
def test_hbm_lattice_chain_beside_lds_chains(use_nlist):
    """A 400x400 chain (160000 lattice bytes, beyond the block walk's LDS) runs
    the scalar walk on its HBM lattice while the small chains of the same batch
    run the block walk in LDS, on a second stream at the same time: every chain
    bit-exact against the oracle for 2 passes, run in two launches."""
    conds = [dict(shape=(400, 400), temperature=2300.0, exchange_potential=0.05, seed=401),
             dict(shape=(24, 24), temperature=1800.0, exchange_potential=-0.1, seed=402),
             dict(shape=(9, 31), temperature=3100.0, exchange_potential=0.2, seed=403)]
    print("case: hbm chain + lds chains use_nlist", use_nlist, flush=True)
    r = cm.Run(conds, mode=cm.MODE_EXACT, J1=0.1, use_nlist=use_nlist)
    r.run(1)
    r.run(1)
    for i, c in enumerate(conds):
        rows, cols = c["shape"]
        o = O.sgc_run(rows, cols, use_nlist=use_nlist, temperature=c["temperature"],
                      mu=c["exchange_potential"], seed=c["seed"], max_count=2)
        compare(r, o, c=i)
    r.close()


@pytest.mark.parametrize("shape,steps", [((12, 10), [1, 37, 119, 120, 5, 240, 3]),
                                          ((64, 64), [4095, 1, 4096, 8191, 2])])
def test_run_steps_inside_passes(shape, steps):
    """mcg_run_steps ends launches inside passes (the block walk cut at any
    attempt): the concatenation equals one uninterrupted run; the position
    (pass, step) and the records at pass ends follow the step counts."""
    rows, cols = shape
    n = rows * cols
    total = sum(steps)
    print("case: run_steps", shape, steps, flush=True)
    a = run_gpu(rows, cols, 2100.0, 0.03, 77, 0)
    for s in steps:
        a.run_steps(s)
    p, st, _ = a.position(0)
    assert (p, st) == (total // n, total % n)
    b = run_gpu(rows, cols, 2100.0, 0.03, 77, total // n)
    b.run_steps(total % n)
    np.testing.assert_array_equal(a.get_occupation(0), b.get_occupation(0))
    np.testing.assert_array_equal(a.rng_dump(0), b.rng_dump(0))
    np.testing.assert_array_equal(a.samples(0), b.samples(0))
    assert a.counters(0) == b.counters(0) and a.position(0) == b.position(0)
    o = O.sgc_run(rows, cols, use_nlist=True, temperature=2100.0, mu=0.03, seed=77,
                  max_count=total // n)
    rec = a.samples(0)
    assert len(rec) == total // n
    np.testing.assert_array_equal(rec["S1"], o["S"])
    # whole passes after a partial one keep the offset into the pass
    a.run(2)
    assert a.position(0)[:2] == (total // n + 2, total % n)
