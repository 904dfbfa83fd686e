"""Benchmark: Metropolis site-attempts/s on one MI355X, batched vs plain.

Default workload (the `value` line): the batched EXACT-mode scan -- 2D square
Ising 128x128, J=0.1 eV, 64 chains over 8 temperatures (500-3000 K) x 8
exchange potentials (-0.4..0.4 eV), seeds 1000+i, all chains in one launch,
bit-exact with the reference's std::mt19937_64 stream.  One step = one pass of
every chain (64 x 16384 site attempts); the timed region runs exactly K steps.

Also measured in the same run and reported beside `value`:
  plain      -- the same EXACT kernel on one chain;
  checker    -- CHECKERBOARD mode: the 64-condition 128x128 scan batched in one
                LDS-resident launch vs the plain one-chain launch run condition
                by condition; and the 4096^2 / 256^3 J1+J2 16-temperature
                families streamed from HBM (batched and one chain);
  fcc        -- the binary FCC cluster expansion: 256 uneven chains (16^3-32^3
                conventional cells) in one autotuned launch vs one chain alone;
  served     -- a Poisson(50/s) stream of 50/50 Ising/FCC run requests through
                the continuous-batching server: throughput and p50/p99 latency;
  cpu_baseline -- the oracle (a restatement of the reference loop with its
                per-step completion check and per-pass sampling) timed on the
                host cores on a bounded sample of the default workload.
"""
import argparse
import concurrent.futures as cf
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LDS_128B_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz = 78.6 TB/s (GB/s units)
HBM_PEAK_GBS = 8000.0
# LDS peak for the rooflines: the 16-byte read rate measured on the device
# (measured_peaks(), set by main()); it exceeds 128 B/clk per CU on MI355X, so
# that figure is no ceiling there.  Until measured: 128 B/clk.
LDS_PEAK = {"gbs": LDS_128B_GBS, "source": "256 CUs x 128 B/clk x 2.4 GHz"}


CLOCK_GHZ = 2.4  # MI355X peak engine clock


def measured_peaks():
    """HBM streaming copy and LDS read bandwidth measured on this device
    (casmcode_monte_amd/lib/libcasm_monte_probes.so)."""
    import ctypes as C
    path = os.path.join(ROOT, "casmcode_monte_amd", "lib", "libcasm_monte_probes.so")
    L = C.CDLL(path)
    L.mcgp_hbm_copy.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
    L.mcgp_lds_read.argtypes = [C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mcgp_last_error.restype = C.c_char_p
    g, t = C.c_double(), C.c_double()
    if L.mcgp_hbm_copy(1 << 30, 20, C.byref(g), C.byref(t)):
        raise RuntimeError(L.mcgp_last_error().decode())
    hbm = dict(gbs=g.value, ms_per_copy=t.value,
               note="1 GiB device-to-device copy, 16-byte loads/stores, read+write bytes")
    if L.mcgp_lds_read(20000, C.byref(g), C.byref(t)):
        raise RuntimeError(L.mcgp_last_error().decode())
    lds = dict(gbs=g.value, ms=t.value,
               note="2 x 1024-thread workgroups per CU reading 16-byte LDS words, no conflicts")
    return dict(hbm_copy=hbm, lds_read=lds, hbm_datasheet_gbs=HBM_PEAK_GBS,
                lds_128B_per_clk_gbs=LDS_128B_GBS)


def set_lds_peak():
    """Measure the LDS read rate once and use it as the LDS roofline peak."""
    pk = measured_peaks()
    LDS_PEAK["gbs"] = pk["lds_read"]["gbs"]
    LDS_PEAK["source"] = "measured 16-byte LDS read rate (measured_peaks.lds_read)"
    return pk


# HBM traffic of one batched exact launch (64 chains of 128x128), FETCH_SIZE x
# 2.0 + WRITE_SIZE x 1.0 of rocprofv3 --pmc at 20 and 2000 passes
# (profiles/exact_block_kernel_64x128sq.md), as fixed + per-pass bytes
EXACT_TRAFFIC_FIXED_B = 2647478.0
EXACT_TRAFFIC_PER_PASS_B = 5120.5


def c2_conditions(shape=(128, 128)):
    out = []
    for a in range(8):
        for b in range(8):
            out.append(dict(shape=shape, temperature=500.0 + a * 2500.0 / 7.0,
                            exchange_potential=-0.4 + b * 0.8 / 7.0, seed=1000 + 8 * a + b))
    return out


def c3_conditions(shape, n=16):
    return [dict(shape=shape, temperature=1000.0 + 200.0 * k, exchange_potential=0.0,
                 seed=0xC3000 + k) for k in range(n)]


def timed_run(run, passes):
    run.reserve_samples(run.n_samples(0) + passes + 1)  # no series growth while timed
    t0 = time.perf_counter()
    run.run(passes)
    wall = time.perf_counter() - t0
    ms, attempts = run.last_run_time()
    return wall, ms / 1e3, attempts


def bench_exact(K, W, extras=True):
    import casmcode_monte_amd as cm
    conds = c2_conditions()
    r = cm.Run(conds, mode=cm.MODE_EXACT, J1=0.1, use_nlist=True)
    r.run(W)
    wall, kern_s, attempts = timed_run(r, K)
    out = dict(wall_s=wall, kernel_s=kern_s, attempts=attempts, chains=len(conds))
    if not extras:
        r.close()
        return out
    # host buffers across PCIe: the lattices in, then every chain's records and
    # final lattice out (the run itself keeps its state in HBM)
    t0 = time.perf_counter()
    for c in range(len(conds)):
        r.set_occupation(c, r.get_occupation(c))
    t1 = time.perf_counter()
    n_rec = 0
    for c in range(len(conds)):
        n_rec += len(r.samples(c))
        r.get_occupation(c)
    t2 = time.perf_counter()
    out["pcie"] = dict(upload_s=t1 - t0, download_s=t2 - t1, records=n_rec,
                       attempts_per_s_pcie_inclusive=attempts / (wall + (t1 - t0) + (t2 - t1)),
                       note="upload timed as a lattice round trip per chain (an upper bound)")
    r.close()
    # plain: one chain, same kernel, fewer passes (rate is per-chain latency bound)
    p = cm.Run(conds[:1], mode=cm.MODE_EXACT, J1=0.1, use_nlist=True)
    p.run(2)
    kp = max(K // 10, 20)
    pw, pk, pa = timed_run(p, kp)
    out["plain"] = dict(passes=kp, wall_s=pw, kernel_s=pk, attempts=pa,
                        attempts_per_s=pa / pw)
    p.close()
    # the 64 chains use 64 of the 256 CUs; 16 independent seeds of each
    # condition (1024 chains, four per CU) show what the chip holds
    reps = [dict(c, seed=c["seed"] + 100000 * k) for k in range(16) for c in conds]
    f = cm.Run(reps, mode=cm.MODE_EXACT, J1=0.1, use_nlist=True)
    f.run(2)
    fp = max(K // 20, 10)
    fw, fk, fa = timed_run(f, fp)
    out["full_chip"] = dict(chains=len(reps), passes=fp, wall_s=fw, kernel_s=fk, attempts=fa,
                            attempts_per_s=fa / fw,
                            note="16 seeds of each of the 64 conditions, four chains per CU")
    f.close()
    return out


def canonical_conditions(shape=(128, 128)):
    """8 temperatures x 8 compositions x = n(+1)/N in 0.1..0.8, random initial
    configurations at that composition (seeded)."""
    import numpy as np
    n = shape[0] * shape[1]
    out, occs = [], []
    for a in range(8):
        for b in range(8):
            out.append(dict(shape=shape, temperature=500.0 + a * 2500.0 / 7.0,
                            seed=2000 + 8 * a + b))
            s = -np.ones(n, dtype=np.int8)
            s[np.random.default_rng(2000 + 8 * a + b).permutation(n)[:round((0.1 + 0.1 * b) * n)]] = 1
            occs.append(s)
    return out, occs


def bench_canonical(passes):
    """Canonical pair-swap chains (CANONICAL mode, bit-exact): 64 chains of 128x128
    in one launch vs one chain alone, and the oracle's rate on one host core."""
    import casmcode_monte_amd as cm
    conds, occs = canonical_conditions()
    r = cm.Run(conds, mode=cm.MODE_CANONICAL, J1=0.1)
    for i, o in enumerate(occs):
        r.set_occupation(i, o)
    r.run(2)
    w, k, a = timed_run(r, passes)
    out = dict(chains=len(conds), passes=passes, wall_s=w, kernel_s=k, attempts=a,
               attempts_per_s=a / w)
    r.close()
    # one 128x128 chain per CU: 4 seeds of each condition fill the chip
    reps = [dict(c, seed=c["seed"] + 100000 * k) for k in range(4) for c in conds]
    f = cm.Run(reps, mode=cm.MODE_CANONICAL, J1=0.1)
    for i in range(len(reps)):
        f.set_occupation(i, occs[i % len(occs)])
    f.run(1)
    fp = max(passes // 4, 10)
    fw, fk, fa = timed_run(f, fp)
    out["full_chip"] = dict(chains=len(reps), passes=fp, wall_s=fw, kernel_s=fk, attempts=fa,
                            attempts_per_s=fa / fw,
                            note="4 seeds of each of the 64 conditions, one chain per CU")
    f.close()
    p = cm.Run(conds[:1], mode=cm.MODE_CANONICAL, J1=0.1)
    p.set_occupation(0, occs[0])
    p.run(1)
    pp = max(passes // 10, 10)
    pw, pk, pa = timed_run(p, pp)
    out["plain_attempts_per_s"] = pa / pw
    out["plain_passes"] = pp
    out["batched_over_plain"] = out["attempts_per_s"] / out["plain_attempts_per_s"]
    p.close()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    cp = 20
    t0 = time.perf_counter()
    O.canonical_run(128, 128, 0.1, conds[0]["temperature"], cp, occs[0], O.mt64_state(2000, 0))
    cw = time.perf_counter() - t0
    out["cpu_port"] = dict(attempts_per_s=cp * 128 * 128 / cw, cores=1,
                           sample=f"1 chain 128x128, {cp} passes, oracle restatement")
    return out


def occupant_conditions(shape=(128, 128)):
    """Two asymmetric units on the two sublattices of the square lattice, unit
    0 allowing (A, B) and unit 1 (B, A); 8 temperatures x 8 exchange potentials
    of B (semi-grand canonical) with random initial occupations (seeded)."""
    import numpy as np
    rows, cols = shape
    n = rows * cols
    units = np.array([((l % rows) + (l // rows)) % 2 for l in range(n)], dtype=np.int8)
    conds, mus, occs = [], [], []
    for a in range(8):
        for b in range(8):
            conds.append(dict(shape=shape, temperature=500.0 + a * 2500.0 / 7.0,
                              seed=3000 + 8 * a + b))
            mus.append([0.0, -0.2 + 0.4 * b / 7.0])
            occs.append(np.random.default_rng(3000 + 8 * a + b).integers(0, 2, n).astype(np.int8))
    return conds, mus, occs, units


def bench_occupant(passes, canonical=False):
    """Occupant events over an OccLocation with two asymmetric units
    (OCCUPANT mode, bit-exact): the 64 conditions in one launch, one chain per
    CU over the whole chip, one chain alone, and the oracle on one host core."""
    import casmcode_monte_amd as cm
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    conds, mus, occs, units = occupant_conditions()
    spec = [[0, 1], [1, 0]]
    swaps = O.occ_canonical_swaps(spec) if canonical else O.occ_semigrand_swaps(spec)

    def make(cs, ms, os_):
        r = cm.Run(cs, mode=cm.MODE_OCCUPANT, J1=0.1,
                   occupant=dict(asym_species=spec, n_species=2, canonical=canonical,
                                 swaps=swaps, exchange_potential=ms))
        for i, o in enumerate(os_):
            r.set_site_asym(i, units)
            r.set_occupation(i, o)
        return r

    r = make(conds, mus, occs)
    r.run(2)
    w, k, a = timed_run(r, passes)
    out = dict(chains=len(conds), passes=passes, wall_s=w, kernel_s=k, attempts=a,
               attempts_per_s=a / w, canonical=canonical,
               system="128x128, units (A,B) and (B,A) on the two sublattices")
    r.close()
    reps = [dict(c, seed=c["seed"] + 100000 * q) for q in range(4) for c in conds]
    f = make(reps, mus * 4, occs * 4)
    f.run(1)
    fp = max(passes // 4, 10)
    fw, fk, fa = timed_run(f, fp)
    out["full_chip"] = dict(chains=len(reps), passes=fp, wall_s=fw, kernel_s=fk, attempts=fa,
                            attempts_per_s=fa / fw,
                            note="4 seeds of each of the 64 conditions, one chain per CU")
    f.close()
    p = make(conds[:1], mus[:1], occs[:1])
    p.run(1)
    pp = max(passes // 10, 10)
    pw, pk, pa = timed_run(p, pp)
    out["plain_attempts_per_s"] = pa / pw
    out["batched_over_plain"] = out["attempts_per_s"] / out["plain_attempts_per_s"]
    p.close()
    cp = 10
    t0 = time.perf_counter()
    O.occ_run(128, 128, units, spec, mus[0], swaps, canonical, 0.1, conds[0]["temperature"],
              cp * 128 * 128, occs[0], O.mt64_state(3000, 0))
    cw = time.perf_counter() - t0
    out["cpu_port"] = dict(attempts_per_s=cp * 128 * 128 / cw, cores=1,
                           sample=f"1 chain 128x128, {cp} passes, oracle restatement")
    return out


def bench_nfold(passes):
    """N-fold way chains (NFOLD mode): the 64 conditions of the exact scan
    (128x128, 8 T x 8 mu) in one launch vs one chain alone, in events/s."""
    import casmcode_monte_amd as cm
    conds = c2_conditions()
    r = cm.Run(conds, mode=cm.MODE_NFOLD, J1=0.1)
    r.run(2)
    w, k, a = timed_run(r, passes)
    out = dict(chains=len(conds), passes=passes, wall_s=w, kernel_s=k, events=a,
               events_per_s=a / w)
    r.close()
    # the chip holds two 128x128 chains per CU: 8 independent replicas (seeds)
    # of each condition fill it
    reps = [dict(c, seed=c["seed"] + 100000 * k) for k in range(8) for c in conds]
    f = cm.Run(reps, mode=cm.MODE_NFOLD, J1=0.1)
    f.run(1)
    fp = max(passes // 8, 5)
    fw, fk, fa = timed_run(f, fp)
    out["full_chip"] = dict(chains=len(reps), passes=fp, wall_s=fw, kernel_s=fk, events=fa,
                            events_per_s=fa / fw,
                            note="8 seeds of each of the 64 conditions, two chains per CU")
    f.close()
    p = cm.Run(conds[:1], mode=cm.MODE_NFOLD, J1=0.1)
    p.run(1)
    pp = max(passes // 10, 5)
    pw, pk, pa = timed_run(p, pp)
    out["plain_events_per_s"] = pa / pw
    out["batched_over_plain"] = out["events_per_s"] / out["plain_events_per_s"]
    p.close()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle as O
    cp = 5
    n = 128 * 128
    tc = time.perf_counter()
    O.nfold_run(128, 128, 0.1, conds[0]["temperature"], conds[0]["exchange_potential"], cp,
                np.ones(n, dtype=np.int8), O.mt64_state(1000, 0))
    cw = time.perf_counter() - tc
    out["cpu_port"] = dict(events_per_s=cp * n / cw, cores=1,
                           sample=f"1 chain 128x128, {cp} passes, oracle restatement")
    return out


def kmc_conditions(shape=(64, 64)):
    """Lattice-gas KMC scan: 8 temperatures (500-3000 K) x 8 atom fractions
    (0.1-0.8), seeds 5000+i, J = 0.05 eV, random initial occupations."""
    import numpy as np
    n = shape[0] * shape[1]
    conds, occs = [], []
    for a in range(8):
        for b in range(8):
            i = 8 * a + b
            conds.append(dict(shape=shape, temperature=500.0 + a * 2500.0 / 7.0, seed=5000 + i))
            occs.append((np.random.default_rng(5000 + i).random(n) < 0.1 * (b + 1))
                        .astype(np.int8))
    return conds, occs


def bench_kmc(passes):
    """Kinetic Monte Carlo with atom tracking (KMC mode): the 64-condition
    lattice-gas scan (64x64, 8 T x 8 atom fractions) in one launch vs one chain
    alone, in events/s (one pass = 4096 events per chain)."""
    import casmcode_monte_amd as cm
    import numpy as np
    conds, occs = kmc_conditions()
    n = 64 * 64

    def make(cs, os_):
        r = cm.Run(cs, mode=cm.MODE_KMC, J1=0.05)
        for c, o in enumerate(os_):
            r.set_occupation(c, o)
        r.run(1)
        return r

    r = make(conds, occs)
    w, k, ev = timed_run(r, passes)
    att = sum(sum(r.counters(c)[1:]) for c in range(len(conds)))
    out = dict(chains=len(conds), passes=passes, wall_s=w, kernel_s=k, events=ev,
               events_per_s=ev / w, attempts_total=att,
               note="attempts_total counts every attempt since the occupation was set "
                    "(one warm-up pass included)")
    r.close()
    reps = [dict(c, seed=c["seed"] + 100000 * j) for j in range(4) for c in conds]
    f = make(reps, occs * 4)
    fp = max(passes // 2, 5)
    fw, fk, fe = timed_run(f, fp)
    out["full_chip"] = dict(chains=len(reps), passes=fp, wall_s=fw, kernel_s=fk, events=fe,
                            events_per_s=fe / fw, note="4 seeds of each of the 64 conditions")
    f.close()
    p = make(conds[:1], occs[:1])
    pp = max(passes // 4, 5)
    pw, pk, pe = timed_run(p, pp)
    out["plain_events_per_s"] = pe / pw
    out["batched_over_plain"] = out["events_per_s"] / out["plain_events_per_s"]
    p.close()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    cp = 10
    tc = time.perf_counter()
    O.kmc_run(64, 64, 0.05, conds[0]["temperature"], cp * n, occs[0], O.mt64_state(5000, 0))
    cw = time.perf_counter() - tc
    out["cpu_port"] = dict(events_per_s=cp * n / cw, cores=1,
                           sample=f"1 chain 64x64, {cp} passes, oracle restatement")
    return out


def bench_checker_scan(passes):
    """64-condition 128^2 scan: batched LDS-resident launch vs one chain at a time."""
    import casmcode_monte_amd as cm
    conds = c2_conditions()
    b = cm.Run(conds, mode=cm.MODE_CHECKERBOARD, J1=0.1, J2=0.0)
    b.run(5)
    sched = b.autotune(20)
    bw, bk, ba = timed_run(b, passes)
    b.close()
    plain_wall, plain_att = 0.0, 0.0
    for c in conds[:8]:  # a bounded sample of the one-chain launches
        p = cm.Run([c], mode=cm.MODE_CHECKERBOARD, J1=0.1, J2=0.0)
        p.run(1)
        w, k, a = timed_run(p, passes)
        plain_wall += w
        plain_att += a
        p.close()
    lds_gbs = 6.0 * ba / bk / 1e9  # own read + 4 neighbour reads + write per attempt
    return dict(passes=passes, batched_attempts_per_s=ba / bw, batched_kernel_s=bk,
                lds_block=sched["lds_block"],
                plain_attempts_per_s=plain_att / plain_wall,
                ratio=(ba / bw) / (plain_att / plain_wall),
                roofline={"bound": "lds", "achieved": lds_gbs, "peak": LDS_PEAK["gbs"],
                          "unit": "GB/s", "frac": lds_gbs / LDS_PEAK["gbs"],
                          "bytes_per_attempt": 6,
                          "kernel": "checker_lds_kernel, 64 chains of 128x128 J1"})


def bench_checker_family(shape, lattice, passes, chains=16):
    import casmcode_monte_amd as cm
    conds = c3_conditions(shape, chains)
    r = cm.Run(conds, mode=cm.MODE_CHECKERBOARD, lattice=lattice, J1=0.1, J2=-0.02,
               sample_period=1)
    r.run(2)
    sched = r.autotune(16)
    w, k, a = timed_run(r, passes)
    r.close()
    return dict(shape=list(shape), chains=chains, passes=passes, wall_s=w, kernel_s=k,
                schedule={k2: sched[k2] for k2 in ("units_per_thread", "graph_passes", "band_lines",
                                                   "line_tiles")},
                attempts=a, attempts_per_s=a / w,
                hbm_gbs_algorithmic=2.0 * a / k / 1e9,
                hbm_frac=2.0 * a / k / 1e9 / HBM_PEAK_GBS)


def bench_fcc(passes, plain=True):
    """Binary FCC cluster expansion: 256 chains of L = 16 + (i mod 17) conventional
    cells (16384-131072 sites), T ~ U(300, 2000) K, mu ~ U(-1, 1) eV, default ECIs,
    all in one autotuned launch; plain = one 32^3 chain alone."""
    import casmcode_monte_amd as cm
    from casmcode_monte_amd import fcc
    conds = fcc.batch_conditions()
    r = cm.Run(conds, mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_FCC, eci=fcc.DEFAULT_ECI)
    r.run(2)
    sched = r.autotune(8)
    w, k, a = timed_run(r, passes)
    r.close()
    if not plain:
        return dict(chains=len(conds), passes=passes, kernel_s=k, attempts_per_s=a / w,
                    lds_block=sched["lds_block"])
    p = cm.Run([conds[16]], mode=cm.MODE_CHECKERBOARD, lattice=cm.LATTICE_FCC,
               eci=fcc.DEFAULT_ECI)
    p.run(2)
    pp = max(passes // 10, 10)
    pw, pk, pa = timed_run(p, pp)
    p.close()
    lds_gbs = 44.0 * a / k / 1e9  # own read + 42 neighbour reads + write per attempt
    return dict(chains=len(conds), passes=passes, wall_s=w, kernel_s=k, attempts=a,
                attempts_per_s=a / w, lds_block=sched["lds_block"],
                plain_attempts_per_s=pa / pw, plain_passes=pp,
                batched_over_plain=(a / w) / (pa / pw),
                roofline={"bound": "lds", "achieved": lds_gbs, "peak": LDS_PEAK["gbs"],
                          "unit": "GB/s", "frac": lds_gbs / LDS_PEAK["gbs"],
                          "bytes_per_attempt": 44})


def bench_served(n_requests):
    """Served scan: a Poisson (50/s) stream of run requests, half 2D Ising (edges
    log-uniform in [16, 512], odd edges with ring colourings) and half FCC cluster expansion
    (edges log-uniform in [8, 48] conventional cells), T in 300-3000 K, mu in
    -1..1 eV, each stopping at <x> precision 1e-3 or 20000 passes."""
    from casmcode_monte_amd import serving
    reqs = serving.poisson_trace(n_requests, rate=50.0, fcc_fraction=0.5)
    srv = serving.Server(fcc_classes=serving.DEFAULT_FCC_CLASSES)
    res, st = srv.serve(reqs)
    srv.close()
    st["converged"] = sum(r.converged for r in res.values())
    st["fcc_requests"] = sum(r.model == "fcc" for r in reqs)
    st["trace"] = (f"first {n_requests} requests of the Poisson(50/s) trace, seed 2024, "
                   "50/50 2D Ising (J1=0.1 eV) / FCC cluster expansion (default ECIs)")
    return st


def cpu_baseline(n_chains, passes):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    conds = c2_conditions()[:n_chains]
    threads = min(n_chains, os.cpu_count() or 1, 16)

    def one(c):
        return O.sgc_run(128, 128, J=0.1, use_nlist=True, temperature=c["temperature"],
                         mu=c["exchange_potential"], seed=c["seed"], max_count=passes,
                         capacity=passes + 4)

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(one, conds))
    wall = time.perf_counter() - t0
    attempts = n_chains * 128 * 128 * passes
    out = dict(value=attempts / wall, unit="site-attempts/s", cores=threads, kind="port",
               sample=f"{n_chains} of the 64 chains x {passes} passes of 128x128 "
                      f"(reference loop with per-step completion check and per-pass "
                      f"sampling), {threads} threads")
    out["configs"] = cpu_configs(O)
    return out


def cpu_configs(O):
    """The reference's CPU path (the oracle's restatement: sequential random-site
    Metropolis) on bounded samples of the other configurations, one core each."""
    import numpy as np
    from casmcode_monte_amd import fcc
    res = {}

    def timed(name, attempts, f, sample):
        t0 = time.perf_counter()
        f()
        w = time.perf_counter() - t0
        res[name] = dict(attempts_per_s=attempts / w, cores=1, sample=sample)

    timed("c1_25x25_end_to_end", 11000 * 625,
          lambda: O.sgc_run(25, 25, J=0.1, use_nlist=True, temperature=2000.0, mu=0.0, seed=42,
                            max_count=11000),
          "25x25, all +1, T=2000 K, seed 42, 11000 passes with per-step completion checks")
    n2 = 4096 * 4096
    timed("c3_4096x4096_J1J2", n2,
          lambda: O.lattice_run(2, (4096, 4096), 0.1, -0.02, 1000.0, 0.0, 0xC3000, 1,
                                np.ones(n2, dtype=np.int8)),
          "one random-site pass of 4096^2 J1+J2 at 1000 K")
    n3 = 256 ** 3
    timed("c3_256cubed_J1J2", n3,
          lambda: O.lattice_run(3, (256, 256, 256), 0.1, -0.02, 1000.0, 0.0, 0xC3000, 1,
                                np.ones(n3, dtype=np.int8)),
          "one random-site pass of 256^3 J1+J2 at 1000 K")
    nf = 4 * 32 ** 3
    timed("c4_fcc_32cubed", 5 * nf,
          lambda: O.fcc_run((32, 32, 32), fcc.DEFAULT_ECI, 1000.0, 0.0, 7, 5,
                            np.zeros(nf, dtype=np.int8)),
          "5 random-site passes of one 32^3 FCC chain at 1000 K")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--skip-extra", action="store_true", help="only the default workload")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--served-requests", type=int, default=400,
                    help="requests of the served scan (0 skips it)")
    ap.add_argument("--only", default="",
                    choices=["", "exact", "scan", "band4096", "sc256", "fcc", "served", "peaks",
                             "occupant", "kmc"],
                    help="run one measurement and print its JSON (for profiling one kernel)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes 10 times "
                         "(the reference point that scales the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_bytes:
        import torch
        a = torch.ones(args.copy_bytes, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes, "launches": 10}))
        return
    if args.gpus != 1:
        print("bench.py: the library runs on one GPU; independent chains shard "
              "trivially across GPUs (each GPU its own handle, no communication) but no "
              "multi-GPU figure is measured",
              file=sys.stderr)
        sys.exit(2)
    import casmcode_monte_amd as cm

    K, W = args.steps, args.warmup
    if args.only:
        import casmcode_monte_amd as cm
        if args.only in ("scan", "fcc"):
            set_lds_peak()
        one = {"exact": lambda: bench_exact(K, W, extras=False),
               "scan": lambda: bench_checker_scan(K),
               "band4096": lambda: bench_checker_family((4096, 4096, 1), cm.LATTICE_SQUARE, K),
               "sc256": lambda: bench_checker_family((256, 256, 256), cm.LATTICE_SC, K),
               "fcc": lambda: bench_fcc(K, plain=False),
               "served": lambda: bench_served(args.served_requests),
               "occupant": lambda: bench_occupant(K),
               "kmc": lambda: bench_kmc(K),
               "peaks": measured_peaks}[args.only]()
        print(json.dumps({"only": args.only, "steps": K, **one}))
        return
    peaks = set_lds_peak()
    ex = bench_exact(K, W)
    print(f"[bench] exact batched done: {ex['attempts'] / ex['wall_s']:.3e} attempts/s",
          file=sys.stderr, flush=True)
    value = ex["attempts"] / ex["wall_s"]
    # The exact walk is one dependent chain per wavefront with the lattice in
    # LDS: neither HBM nor MFMA bounds it, the latency of each step does.  The
    # roofline figure is the algorithmic LDS traffic (own read + 4 neighbour
    # reads + write = 6 B per attempt) over the measured kernel time.
    bytes_per_attempt = 6.0
    lds_achieved = bytes_per_attempt * ex["attempts"] / ex["kernel_s"] / 1e9
    attempts_per_chain = ex["attempts"] / ex["chains"]
    cycles_per_attempt = ex["kernel_s"] * CLOCK_GHZ * 1e9 / attempts_per_chain
    # HBM bytes one launch must move: per chain the int8 lattice in and out,
    # the engine/counter state (2600 B) in and out, the parameters (~220 B) in,
    # and one 80 B record per pass; measured traffic from the fitted model
    n_sites = 128 * 128
    hbm_bytes = ex["chains"] * (2 * n_sites + 2 * 2600 + 220 + 80 * K)
    traffic = EXACT_TRAFFIC_FIXED_B + EXACT_TRAFFIC_PER_PASS_B * K
    line = {
        "metric": "Metropolis site-attempts/s, 1 MI355X, batched vs plain; p50/p99 run latency",
        "value": value,
        "unit": "site-attempts/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ex["wall_s"] * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8+u64+f64",
        "data": "synthetic (all +1 initial lattices, seeded MT19937-64 streams)",
        "config": {"workload": "batched exact mode: 2D Ising 128x128, J=0.1 eV, 64 chains "
                               "(8 T in 500-3000 K x 8 mu in -0.4..0.4 eV), seeds 1000+i, "
                               "MT19937-64, FP64 running energy, one chain per wavefront",
                   "chains": 64, "sites_per_chain": 16384, "passes": K,
                   "parallelism": "1 GPU, 64 workgroups (one 64-lane wavefront per chain)"},
        "roofline": {"bound": "latency", "achieved": lds_achieved, "peak": LDS_PEAK["gbs"],
                     "unit": "GB/s", "frac": lds_achieved / LDS_PEAK["gbs"],
                     "peak_source": LDS_PEAK["source"],
                     "traffic": traffic,
                     "kernel": "exact_block_kernel",
                     "bytes_per_attempt": 6,
                     "cycles_per_attempt_per_chain": cycles_per_attempt,
                     "note": "a dependent walk per chain with the lattice in LDS: the latency "
                             "of each step binds it; achieved = 6 B of LDS per attempt over "
                             "the kernel time; traffic = HBM bytes per launch (rocprofv3 "
                             "FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, fitted fixed + per-pass)",
                     "hbm": {"algorithmic_bytes_per_launch": hbm_bytes,
                             "achieved_gbs": hbm_bytes / ex["kernel_s"] / 1e9,
                             "peak": HBM_PEAK_GBS,
                             "traffic_over_algorithmic": traffic / hbm_bytes}},
        "latency_s": {"p50": ex["wall_s"], "p99": ex["wall_s"],
                      "note": "all 64 runs complete in the same launch"},
        "plain": {"attempts_per_s": ex["plain"]["attempts_per_s"],
                  "batched_over_plain": value / ex["plain"]["attempts_per_s"],
                  "passes": ex["plain"]["passes"]},
        "exact_full_chip": ex["full_chip"],
        "pcie_inclusive": ex["pcie"],
    }
    if not args.skip_extra:
        scan = bench_checker_scan(200)
        print(f"[bench] checker scan done: ratio {scan['ratio']:.1f}", file=sys.stderr,
              flush=True)
        fam2 = bench_checker_family((4096, 4096, 1), cm.LATTICE_SQUARE, 50)
        print("[bench] checker 4096^2 done", file=sys.stderr, flush=True)
        fam2_plain = bench_checker_family((4096, 4096, 1), cm.LATTICE_SQUARE, 50, chains=1)
        fam3 = bench_checker_family((256, 256, 256), cm.LATTICE_SC, 20)
        print("[bench] checker 256^3 done", file=sys.stderr, flush=True)
        fc = bench_fcc(200)
        print(f"[bench] fcc batch done: {fc['attempts_per_s']:.3e} attempts/s", file=sys.stderr,
              flush=True)
        line["fcc_cluster_expansion_256_chains"] = fc
        cn = bench_canonical(max(K // 4, 20))
        print(f"[bench] canonical batch done: {cn['attempts_per_s']:.3e} attempts/s",
              file=sys.stderr, flush=True)
        line["canonical_exact_64_chains"] = cn
        nf = bench_nfold(max(K // 10, 20))
        print(f"[bench] nfold batch done: {nf['events_per_s']:.3e} events/s", file=sys.stderr,
              flush=True)
        line["nfold_64_chains"] = nf
        oc = bench_occupant(max(K // 10, 20))
        print(f"[bench] occupant batch done: {oc['attempts_per_s']:.3e} attempts/s",
              file=sys.stderr, flush=True)
        line["occupant_sgc_two_units_64_chains"] = oc
        km = bench_kmc(max(K // 20, 20))
        print(f"[bench] kmc batch done: {km['events_per_s']:.3e} events/s", file=sys.stderr,
              flush=True)
        line["kmc_lattice_gas_64_chains"] = km
        line["checkerboard"] = {
            "scan_128x128_64_conditions": scan,
            "square_4096x4096_J1J2_16T": fam2,
            "square_4096x4096_J1J2_1chain": fam2_plain,
            "sc_256cubed_J1J2_16T": fam3,
        }
        if args.served_requests > 0:
            served = bench_served(args.served_requests)
            print(f"[bench] served done: {served['attempts_per_s']:.3e} attempts/s",
                  file=sys.stderr, flush=True)
            line["served"] = served
            line["latency_s"] = {"p50": served["p50_latency_s"],
                                 "p99": served["p99_latency_s"],
                                 "note": "served scan run latency (arrival to result)",
                                 "exact_batch_wall_s": ex["wall_s"]}
        line["roofline_checkerboard"] = {
            "bound": "hbm", "achieved": fam2["hbm_gbs_algorithmic"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": fam2["hbm_frac"], "traffic": None,
            "kernel": ("checker_band_kernel + checker_seam_kernel" if fam2["schedule"].get("band_lines")
                       else "checker_color_kernel") + ", 16 x 4096^2 J1+J2 (2 B/attempt algorithmic)"}
        line["roofline_checkerboard_3d"] = {
            "bound": "hbm", "achieved": fam3["hbm_gbs_algorithmic"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": fam3["hbm_frac"], "traffic": None,
            "kernel": ("checker_slab_kernel + checker_slab_seam_kernel"
                       if fam3["schedule"].get("band_lines") else "checker_pair_kernel")
                      + ", 16 x 256^3 J1+J2 (2 B/attempt algorithmic)"}
    if not args.skip_extra:
        line["measured_peaks"] = peaks
    if not args.skip_cpu:
        line["cpu_baseline"] = cpu_baseline(16, 1000)
        print("[bench] cpu baseline done", file=sys.stderr, flush=True)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
