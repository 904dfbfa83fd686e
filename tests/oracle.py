"""ctypes bindings of the oracle (oracle/lib/liboracle.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only.  The oracle restates the reference algorithm on the CPU
(see oracle/refcpu.hh for the reference file:line each function follows).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "lib", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class SgcParams(C.Structure):
    _fields_ = [
        ("rows", C.c_int64), ("cols", C.c_int64), ("J", C.c_double),
        ("use_nlist", C.c_int64), ("temperature", C.c_double), ("mu", C.c_double),
        ("seed", C.c_uint64), ("sample_period", C.c_int64),
        ("init_occ", C.POINTER(C.c_int32)),
        ("min_count", C.c_int64), ("min_sample", C.c_int64),
        ("max_count", C.c_int64), ("max_sample", C.c_int64),
        ("check_begin", C.c_int64), ("check_period", C.c_int64), ("log_spacing", C.c_int64),
        ("check_base", C.c_double), ("check_shift", C.c_double),
        ("check_period_max", C.c_int64),
        ("prec_formation_energy", C.c_double), ("prec_param_composition", C.c_double),
        ("prec_potential_energy", C.c_double),
        ("trace_len", C.c_int64),
    ]


class SgcOutput(C.Structure):
    _fields_ = [
        ("capacity", C.c_int64),
        ("formation_energy", C.POINTER(C.c_double)),
        ("param_composition", C.POINTER(C.c_double)),
        ("potential_energy", C.POINTER(C.c_double)),
        ("S", C.POINTER(C.c_int64)), ("M", C.POINTER(C.c_int64)),
        ("final_occ", C.POINTER(C.c_int32)),
        ("trace", C.POINTER(C.c_int64)),
        ("engine", C.POINTER(C.c_uint64)),
        ("n_samples", C.c_int64), ("n_pass", C.c_int64), ("n_accept", C.c_int64),
        ("n_reject", C.c_int64), ("n_checks", C.c_int64),
        ("is_complete", C.c_int64), ("all_equilibrated", C.c_int64), ("N_eq", C.c_int64),
        ("all_converged", C.c_int64), ("N_stats", C.c_int64),
        ("mean", C.c_double * 3), ("prec", C.c_double * 3), ("converged", C.c_int64 * 3),
    ]


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _declare(L):
    L.orc_sgc_run.argtypes = [C.POINTER(SgcParams), C.POINTER(SgcOutput)]
    L.orc_sgc_run.restype = C.c_int
    L.orc_mt64_outputs.argtypes = [C.c_uint64, C.c_int64, C.POINTER(C.c_uint64)]
    L.orc_mt64_state.argtypes = [C.c_uint64, C.c_int64, C.POINTER(C.c_uint64)]
    L.orc_uniform_int_replay.argtypes = [C.POINTER(C.c_uint64), C.c_int64, C.c_long,
                                         C.POINTER(C.c_int64)]
    L.orc_uniform_int_replay.restype = C.c_long
    L.orc_stochastic_count_steps.argtypes = [C.c_uint64, C.c_double, C.c_int64,
                                             C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]
    L.orc_stochastic_count_steps.restype = None
    L.orc_canonical.argtypes = [C.c_uint64]
    L.orc_canonical.restype = C.c_double
    L.orc_accept_prob.argtypes = [C.c_double, C.c_double]
    L.orc_accept_prob.restype = C.c_double
    L.orc_kb.restype = C.c_double
    L.orc_basic_statistics.argtypes = [C.POINTER(C.c_double), C.c_int64, C.c_double,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_weighted_statistics.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.c_int64, C.c_double, C.c_int64, C.c_int64,
                                          C.POINTER(C.c_double)]
    L.orc_weighted_statistics.restype = C.c_int
    L.orc_resample.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                               C.c_double, C.c_int64, C.POINTER(C.c_double)]
    L.orc_resample.restype = None
    L.orc_autocorrelation_factor.argtypes = [C.POINTER(C.c_double), C.c_int64]
    L.orc_autocorrelation_factor.restype = C.c_double
    L.orc_autocorrelation_factor_inc.argtypes = [C.POINTER(C.c_double), C.c_int64, C.c_double]
    L.orc_autocorrelation_factor_inc.restype = C.c_double
    L.orc_completion_check_replay.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double,
                                              C.c_double, C.c_double, C.c_double, C.c_double,
                                              C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.orc_completion_check_replay.restype = None
    L.orc_equilibration.argtypes = [C.POINTER(C.c_double), C.c_int64, C.c_double,
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.orc_approx_erf_inv.argtypes = [C.c_double]
    L.orc_approx_erf_inv.restype = C.c_double
    L.orc_ising_energy.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_int32), C.c_double,
                                   C.c_int64]
    L.orc_ising_energy.restype = C.c_double
    L.orc_ising_delta.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_int32), C.c_double,
                                  C.c_int64, C.c_double, C.c_int64, C.c_int32]
    L.orc_ising_delta.restype = C.c_double
    L.orc_sgc_potential.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_int32), C.c_double,
                                    C.c_int64, C.c_double]
    L.orc_sgc_potential.restype = C.c_double
    L.orc_lattice_run.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_uint64, C.c_int64,
                                  C.POINTER(C.c_int8), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.orc_lattice_observables.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int8),
                                          C.POINTER(C.c_int64)]
    L.orc_lattice_enumerate.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.c_double,
                                        C.c_double, C.c_double, C.c_double,
                                        C.POINTER(C.c_double)]
    P64, P8, PD = C.POINTER(C.c_int64), C.POINTER(C.c_int8), C.POINTER(C.c_double)
    L.orc_fcc_counts.argtypes = [P64, P8, P64]
    L.orc_fcc_local.argtypes = [P64, P8, C.c_int64, P64]
    L.orc_fcc_shell_sizes.argtypes = [P64]
    L.orc_fcc_energy.argtypes = [P64, PD, P8]
    L.orc_fcc_energy.restype = C.c_double
    L.orc_fcc_run.argtypes = [P64, PD, C.c_double, C.c_double, C.c_uint64, C.c_int64, P8, P64]
    PU = C.POINTER(C.c_uint64)
    L.orc_canonical_run.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int64,
                                    P8, PU, P64, P64, P64, PD, P64, P64, PD, C.c_int64,
                                    C.c_char_p]
    L.orc_canonical_run.restype = C.c_int
    L.orc_occ_run.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                              C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                              C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                              C.c_int64, C.c_char_p]
    L.orc_occ_run.restype = C.c_int
    L.orc_kmc_run.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int64,
                              C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_kmc_run.restype = C.c_int64
    L.orc_nfold_run.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double,
                                C.c_int64, C.c_int64, C.c_double, C.c_int64, P8, PU, P64, P64,
                                PD, PD, P64, PD, P64, C.c_int64, PD]
    L.orc_nfold_run.restype = C.c_int64


KB = None


def kb():
    return lib().orc_kb()


def mt64_outputs(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_mt64_outputs(seed, n, _p(out, C.c_uint64))
    return out


def mt64_state(seed, n_discard):
    out = np.zeros(313, dtype=np.uint64)
    lib().orc_mt64_state(seed, n_discard, _p(out, C.c_uint64))
    return out


def uniform_int_replay(words, maximum):
    w = np.ascontiguousarray(words, dtype=np.uint64)
    used = C.c_int64(0)
    v = lib().orc_uniform_int_replay(_p(w, C.c_uint64), len(w), maximum, C.byref(used))
    return v, used.value


def canonical(x):
    return lib().orc_canonical(int(x))


def accept_prob(dE, T):
    return lib().orc_accept_prob(dE, T)


def basic_statistics(x, confidence=0.95):
    x = np.ascontiguousarray(x, dtype=np.float64)
    m, p = C.c_double(), C.c_double()
    lib().orc_basic_statistics(_p(x, C.c_double), len(x), confidence, C.byref(m), C.byref(p))
    return m.value, p.value


def weighted_statistics(x, w, confidence=0.95, method=1, n_resamples=10000):
    """BasicStatisticsCalculator on weighted observations (BasicStatistics.cc:144-188):
    dict of mean, calculated_precision, variance, autocorr_factor."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    if len(x) != len(w):
        raise ValueError("observations.size() != sample_weight.size()")
    out = np.zeros(4)
    if lib().orc_weighted_statistics(_p(x, C.c_double), _p(w, C.c_double), len(x), confidence,
                                     method, n_resamples, _p(out, C.c_double)):
        raise ValueError("weighted statistics: empty or mismatched input, or bad method")
    return dict(mean=out[0], calculated_precision=out[1], variance=out[2], autocorr_factor=out[3])


def resample(x, w, w_sum, m):
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.zeros(m)
    lib().orc_resample(_p(x, C.c_double), _p(w, C.c_double), len(x), w_sum, m,
                       _p(out, C.c_double))
    return out


def stochastic_count_steps(seed, rate, n):
    """n successive stochastic_count_step(rate) draws (SamplingParams.hh:193-204)
    from a fresh engine, and the engine's next output after them."""
    dn = np.zeros(n, dtype=np.int64)
    nxt = C.c_uint64(0)
    lib().orc_stochastic_count_steps(seed, rate, n, _p(dn, C.c_int64), C.byref(nxt))
    return dn, nxt.value


def autocorrelation_factor(x, increment=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    if increment is None:
        return lib().orc_autocorrelation_factor(_p(x, C.c_double), len(x))
    return lib().orc_autocorrelation_factor_inc(_p(x, C.c_double), len(x), increment)


def completion_check_replay(seed, min_sample=100, e_mean=1.0, e_amp=0.1, e_abs=0.001,
                            v_mean=20.0, v_amp=1.0, v_abs=0.01):
    """The loop of the reference's python/tests/sampling/test_CompletionCheck.py:40-124
    on the oracle's CompletionCheck, with a seeded engine."""
    oi = np.zeros(8, dtype=np.int64)
    od = np.zeros(4)
    lib().orc_completion_check_replay(seed, min_sample, e_mean, e_amp, e_abs, v_mean, v_amp,
                                      v_abs, _p(oi, C.c_int64), _p(od, C.c_double))
    keys = ("n_steps", "n_samples", "is_complete", "all_equilibrated", "n_equilibration_results",
            "all_converged", "n_convergence_results", "N_samples_for_statistics")
    out = {k: int(v) for k, v in zip(keys, oi)}
    out.update(prec_e=od[0], prec_v=od[1], mean_e=od[2], mean_v=od[3])
    return out


def equilibration(x, prec):
    x = np.ascontiguousarray(x, dtype=np.float64)
    a, b = C.c_int64(), C.c_int64()
    lib().orc_equilibration(_p(x, C.c_double), len(x), prec, C.byref(a), C.byref(b))
    return bool(a.value), b.value


def ising_energy(occ2d_colmajor, rows, cols, J, use_nlist):
    o = np.ascontiguousarray(occ2d_colmajor, dtype=np.int32)
    return lib().orc_ising_energy(rows, cols, _p(o, C.c_int32), J, int(use_nlist))


def ising_delta(occ, rows, cols, J, use_nlist, mu, l, new_occ):
    o = np.ascontiguousarray(occ, dtype=np.int32)
    return lib().orc_ising_delta(rows, cols, _p(o, C.c_int32), J, int(use_nlist), mu, l, new_occ)


def sgc_potential(occ, rows, cols, J, use_nlist, mu):
    o = np.ascontiguousarray(occ, dtype=np.int32)
    return lib().orc_sgc_potential(rows, cols, _p(o, C.c_int32), J, int(use_nlist), mu)


def sgc_run(rows, cols, J=0.1, use_nlist=True, temperature=2000.0, mu=0.0, seed=42,
            sample_period=1, init_occ=None, min_count=-1, min_sample=-1, max_count=-1,
            max_sample=-1, check_begin=100, check_period=100, log_spacing=False,
            check_base=10.0, check_shift=2.0, check_period_max=10000,
            prec=(0.0, 0.0, 0.0), trace_len=0, capacity=None):
    """Run the restated SGC Ising loop; returns a dict of numpy arrays/scalars."""
    p = SgcParams()
    p.rows, p.cols, p.J, p.use_nlist = rows, cols, J, int(use_nlist)
    p.temperature, p.mu, p.seed, p.sample_period = temperature, mu, seed, sample_period
    occ_keep = None
    if init_occ is not None:
        occ_keep = np.ascontiguousarray(init_occ, dtype=np.int32)
        p.init_occ = _p(occ_keep, C.c_int32)
    p.min_count, p.min_sample, p.max_count, p.max_sample = min_count, min_sample, max_count, max_sample
    p.check_begin, p.check_period, p.log_spacing = check_begin, check_period, int(log_spacing)
    p.check_base, p.check_shift, p.check_period_max = check_base, check_shift, check_period_max
    p.prec_formation_energy, p.prec_param_composition, p.prec_potential_energy = prec
    p.trace_len = trace_len
    if capacity is None:
        capacity = (max_count if max_count > 0 else 100000) // max(sample_period, 1) + 10
    fe = np.zeros(capacity); pc = np.zeros(capacity); pe = np.zeros(capacity)
    S = np.zeros(capacity, dtype=np.int64); M = np.zeros(capacity, dtype=np.int64)
    fin = np.zeros(rows * cols, dtype=np.int32)
    tr = np.zeros(max(trace_len, 1) * 5, dtype=np.int64)
    eng = np.zeros(313, dtype=np.uint64)
    o = SgcOutput()
    o.capacity = capacity
    o.formation_energy, o.param_composition = _p(fe, C.c_double), _p(pc, C.c_double)
    o.potential_energy = _p(pe, C.c_double)
    o.S, o.M = _p(S, C.c_int64), _p(M, C.c_int64)
    o.final_occ, o.trace, o.engine = _p(fin, C.c_int32), _p(tr, C.c_int64), _p(eng, C.c_uint64)
    rc = lib().orc_sgc_run(C.byref(p), C.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_sgc_run failed rc={rc} (n_samples={o.n_samples})")
    n = o.n_samples
    return dict(
        formation_energy=fe[:n], param_composition=pc[:n], potential_energy=pe[:n],
        S=S[:n], M=M[:n], final_occ=fin, trace=tr[: trace_len * 5].reshape(-1, 5),
        engine=eng, n_samples=n, n_pass=o.n_pass, n_accept=o.n_accept, n_reject=o.n_reject,
        n_checks=o.n_checks, is_complete=bool(o.is_complete),
        all_equilibrated=bool(o.all_equilibrated), N_eq=o.N_eq,
        all_converged=bool(o.all_converged), N_stats=o.N_stats,
        mean=list(o.mean), prec=list(o.prec), converged=list(o.converged),
    )


def lattice_run(ndim, L, J1, J2, T, mu, seed, n_passes, occ):
    Ls = np.array(list(L) + [1] * (3 - len(L)), dtype=np.int64)
    o = np.ascontiguousarray(occ, dtype=np.int8).copy()
    S1 = np.zeros(n_passes, dtype=np.int64); S2 = np.zeros(n_passes, dtype=np.int64)
    M = np.zeros(n_passes, dtype=np.int64)
    lib().orc_lattice_run(ndim, _p(Ls, C.c_int64), J1, J2, T, mu, seed, n_passes,
                          _p(o, C.c_int8), _p(S1, C.c_int64), _p(S2, C.c_int64), _p(M, C.c_int64))
    return o, S1, S2, M


def lattice_observables(ndim, L, occ):
    Ls = np.array(list(L) + [1] * (3 - len(L)), dtype=np.int64)
    o = np.ascontiguousarray(occ, dtype=np.int8)
    out = np.zeros(3, dtype=np.int64)
    lib().orc_lattice_observables(ndim, _p(Ls, C.c_int64), _p(o, C.c_int8), _p(out, C.c_int64))
    return tuple(int(v) for v in out)


def lattice_enumerate(ndim, L, J1, J2, T, mu):
    Ls = np.array(list(L) + [1] * (3 - len(L)), dtype=np.int64)
    out = np.zeros(6)
    rc = lib().orc_lattice_enumerate(ndim, _p(Ls, C.c_int64), J1, J2, T, mu, _p(out, C.c_double))
    if rc:
        raise ValueError("lattice too large to enumerate")
    return dict(E=out[0], E2=out[1], Nx=out[2], Nx2=out[3], ENx=out[4], Ef=out[5])


# ---- binary FCC cluster expansion ------------------------------------------------
FCC_FIELDS = ("N1", "P1", "P2", "P3", "T")


def fcc_counts(L, occ):
    Ls = np.array(L, dtype=np.int64)
    o = np.ascontiguousarray(occ, dtype=np.int8)
    out = np.zeros(5, dtype=np.int64)
    lib().orc_fcc_counts(_p(Ls, C.c_int64), _p(o, C.c_int8), _p(out, C.c_int64))
    return dict(zip(FCC_FIELDS, (int(v) for v in out)))


def fcc_local(L, occ, l):
    Ls = np.array(L, dtype=np.int64)
    o = np.ascontiguousarray(occ, dtype=np.int8)
    out = np.zeros(4, dtype=np.int64)
    lib().orc_fcc_local(_p(Ls, C.c_int64), _p(o, C.c_int8), l, _p(out, C.c_int64))
    return tuple(int(v) for v in out)


def fcc_shell_sizes():
    out = np.zeros(4, dtype=np.int64)
    lib().orc_fcc_shell_sizes(_p(out, C.c_int64))
    return tuple(int(v) for v in out)


def fcc_energy(L, eci, occ):
    Ls = np.array(L, dtype=np.int64)
    e = np.ascontiguousarray(eci, dtype=np.float64)
    o = np.ascontiguousarray(occ, dtype=np.int8)
    return lib().orc_fcc_energy(_p(Ls, C.c_int64), _p(e, C.c_double), _p(o, C.c_int8))


def fcc_run(L, eci, T, mu, seed, n_passes, occ):
    """Sequential random-site SGC Metropolis; returns (final occ, counts[n_passes, 5])."""
    Ls = np.array(L, dtype=np.int64)
    e = np.ascontiguousarray(eci, dtype=np.float64)
    o = np.ascontiguousarray(occ, dtype=np.int8).copy()
    out = np.zeros((n_passes, 5), dtype=np.int64)
    lib().orc_fcc_run(_p(Ls, C.c_int64), _p(e, C.c_double), T, mu, seed, n_passes,
                      _p(o, C.c_int8), _p(out, C.c_int64))
    return o, out


def fcc_coords(L):
    """Doubled coordinates (x, y, z) of every site in the linear order
    l = h + L0*(y + 2*L1*z), x = ((y+z)&1) + 2h."""
    L0, L1, L2 = L
    l = np.arange(4 * L0 * L1 * L2)
    h = l % L0
    line = l // L0
    y = line % (2 * L1)
    z = line // (2 * L1)
    x = ((y + z) & 1) + 2 * h
    return x, y, z



class CanonicalError(RuntimeError):
    """The reference's canonical loop would stop with this exception."""


def canonical_run(rows, cols, J, T, n_passes, occ, engine, loc=None, use_nlist=True):
    """Canonical pair-swap Metropolis (occupation_metropolis + propose_canonical_event
    over an OccLocation) on spins occ (+1/-1, column-major).

    engine: 313 words (std::mt19937_64 state, position), e.g. mt64_state(seed, 0).
    loc: None (lists built by OccLocation::initialize) or the species-0 list then
    the species-1 list.  use_nlist: variant of the initial per_supercell energy.  Returns a dict with e_run and S per pass, final occ,
    engine, loc (lists), mol_loc, n_accept, n_reject and the initial energy.
    """
    n = rows * cols
    o = np.ascontiguousarray(occ, dtype=np.int8).copy()
    eng = np.ascontiguousarray(engine, dtype=np.uint64).copy()
    li = None if loc is None else np.ascontiguousarray(loc, dtype=np.int64)
    lo = np.zeros(n, dtype=np.int64)
    ml = np.zeros(n, dtype=np.int64)
    e = np.zeros(max(n_passes, 1), dtype=np.float64)
    S = np.zeros(max(n_passes, 1), dtype=np.int64)
    cnt = np.zeros(2, dtype=np.int64)
    init = C.c_double(0.0)
    err = C.create_string_buffer(256)
    rc = lib().orc_canonical_run(rows, cols, J, T, n_passes, _p(o, C.c_int8),
                                 _p(eng, C.c_uint64), None if li is None else _p(li, C.c_int64),
                                 _p(lo, C.c_int64), _p(ml, C.c_int64), _p(e, C.c_double),
                                 _p(S, C.c_int64), _p(cnt, C.c_int64), C.byref(init),
                                 int(use_nlist), err)
    if rc:
        raise CanonicalError(err.value.decode())
    return dict(e_run=e[:n_passes], S=S[:n_passes], occ=o, engine=eng, loc=lo, mol_loc=ml,
                n_accept=int(cnt[0]), n_reject=int(cnt[1]), initial=init.value)


class OccError(RuntimeError):
    """The reference's occupant-event loop would stop with this exception."""


def occ_run(rows, cols, site_asym, asym_species, mu, swaps, canonical, J, T, n_steps, occ,
            engine, loc=None, n_unitcells=None, use_nlist=True, cap=None):
    """Occupant events over an OccLocation with several asymmetric units
    (occupation_metropolis with propose_semigrand_canonical_event, or with
    propose_canonical_event when canonical) on occupant indices occ
    (column-major).

    site_asym: unit of each site; asym_species: per unit, the species index of
    each occupant (1 or 2 entries); mu: exchange potential per species; swaps:
    (cand_a, cand_b) pairs of candidate indices.  loc: None (lists built by
    OccLocation::initialize) or the candidate lists' mol ids in candidate order
    (mol ids number the mutating sites in site order).  Returns a dict with
    e_run, S, M per completed pass, final occ, engine, loc (mol ids), loc_sites
    (the same lists as site indices), n_accept, n_reject, n_steps, initial and
    error (the exception text, or None).
    """
    n = rows * cols
    sa = np.ascontiguousarray(site_asym, dtype=np.int8)
    n_asym = len(asym_species)
    spec = np.zeros(2 * n_asym, dtype=np.int64)
    osz = np.zeros(n_asym, dtype=np.int64)
    for b, sp in enumerate(asym_species):
        osz[b] = len(sp)
        spec[2 * b:2 * b + len(sp)] = sp
    m = np.ascontiguousarray(mu, dtype=np.float64)
    sw = np.ascontiguousarray(np.asarray(swaps, dtype=np.int64).reshape(-1))
    o = np.ascontiguousarray(occ, dtype=np.int8).copy()
    eng = np.ascontiguousarray(engine, dtype=np.uint64).copy()
    li = None if loc is None else np.ascontiguousarray(loc, dtype=np.int64)
    lo = np.zeros(n, dtype=np.int64)
    n_mut = int(sum(osz[b] > 1 for b in sa))
    if cap is None:
        cap = n_steps // max(1, n_mut) + 2
    e = np.zeros(max(cap, 1), dtype=np.float64)
    S = np.zeros(max(cap, 1), dtype=np.int64)
    M = np.zeros(max(cap, 1), dtype=np.int64)
    cnt = np.zeros(4, dtype=np.int64)
    init = C.c_double(0.0)
    err = C.create_string_buffer(256)
    rc = lib().orc_occ_run(rows, cols, _p(sa, C.c_int8), n_asym, _p(spec, C.c_int64),
                           _p(osz, C.c_int64), len(m), _p(m, C.c_double), _p(sw, C.c_int64),
                           len(sw) // 2, int(bool(canonical)), J, T,
                           float(n if n_unitcells is None else n_unitcells), n_steps,
                           _p(o, C.c_int8), _p(eng, C.c_uint64),
                           None if li is None else _p(li, C.c_int64), _p(lo, C.c_int64),
                           _p(e, C.c_double), _p(S, C.c_int64), _p(M, C.c_int64), cap,
                           _p(cnt, C.c_int64), C.byref(init), int(use_nlist), err)
    np_ = min(int(cnt[3]), cap)
    mol_site = np.array([l for l in range(n) if osz[sa[l]] > 1], dtype=np.int64)
    out = dict(e_run=e[:np_], S=S[:np_], M=M[:np_], occ=o, engine=eng, loc=lo[:n_mut],
               loc_sites=mol_site[lo[:n_mut]],
               n_accept=int(cnt[0]), n_reject=int(cnt[1]), n_steps=int(cnt[2]),
               initial=init.value, error=None)
    if rc:
        msg = err.value.decode()
        if int(cnt[2]) == 0 and not msg.startswith("Error in choose_"):
            raise OccError(msg)
        out["error"] = msg
    return out


def occ_candidates(asym_species):
    """OccCandidateList order (OccCandidate.cc:31-52): (unit, species) of every
    unit with more than one occupant, in occupant order."""
    return [(b, s) for b, sp in enumerate(asym_species) if len(sp) > 1 for s in sp]


def occ_semigrand_swaps(asym_species):
    """make_semigrand_canonical_swaps (OccCandidate.cc:129-143) as candidate pairs."""
    c = occ_candidates(asym_species)
    return [(i, j) for i in range(len(c)) for j in range(len(c))
            if c[i][0] == c[j][0] and c[i][1] != c[j][1]]


def occ_canonical_swaps(asym_species):
    """make_canonical_swaps (OccCandidate.cc:99-110) as candidate pairs."""
    c = occ_candidates(asym_species)
    allowed = lambda b, s: s in asym_species[b]
    return [(i, j) for i in range(len(c)) for j in range(len(c))
            if c[i] < c[j] and c[i][1] != c[j][1] and allowed(c[i][0], c[j][1])
            and allowed(c[j][0], c[i][1])]



def kmc_run(rows, cols, J, T, n_events, occ, engine, time_period=0.0, cap=None):
    """Lattice-gas kinetic Monte Carlo (kinetic_monte_carlo.hh loop with the
    rejection selector of refcpu::kmc_run) on occupant indices occ (0 vacancy,
    1 atom).  time_period > 0 samples on the grid k * time_period, otherwise
    after every pass (rows*cols events).  Returns a dict with the records
    (events, time, S, sum_sq, sum_dx, sum_dy), final occ, engine, loc (vacancy
    list then atom list), atoms (site, dx, dy, n_jumps per atom), time and
    attempts."""
    n = rows * cols
    o = np.ascontiguousarray(occ, dtype=np.int8).copy()
    eng = np.ascontiguousarray(engine, dtype=np.uint64).copy()
    n_atoms = int(o.sum())
    if cap is None:
        cap = 1 << 16
    lo = np.zeros(n, dtype=np.int64)
    at = np.zeros(4 * max(n_atoms, 1), dtype=np.int64)
    ev = np.zeros(cap, dtype=np.int64)
    tm = np.zeros(cap, dtype=np.float64)
    S = np.zeros(cap, dtype=np.int64)
    sq = np.zeros(cap, dtype=np.int64)
    dx = np.zeros(cap, dtype=np.int64)
    dy = np.zeros(cap, dtype=np.int64)
    ft = C.c_double(0.0)
    att = C.c_int64(0)
    nr = lib().orc_kmc_run(rows, cols, J, T, n_events, time_period, _p(o, C.c_int8),
                           _p(eng, C.c_uint64), _p(lo, C.c_int64), _p(at, C.c_int64), cap,
                           _p(ev, C.c_int64), _p(tm, C.c_double), _p(S, C.c_int64),
                           _p(sq, C.c_int64), _p(dx, C.c_int64), _p(dy, C.c_int64),
                           C.byref(ft), C.byref(att))
    k = min(nr, cap)
    return dict(events=ev[:k], time=tm[:k], S=S[:k], sum_sq=sq[:k], sum_dx=dx[:k],
                sum_dy=dy[:k], n_records=nr, occ=o, engine=eng, loc=lo,
                atoms=at[:4 * n_atoms].reshape(n_atoms, 4), final_time=ft.value,
                attempts=att.value)

def nfold_run(rows, cols, J, T, mu, n_passes, occ, engine, sample_period=1, time_period=0.0,
              capacity=None, sample_start=False):
    """N-fold way (class-based rejection-free selector) on spins occ, sampled by
    pass (time_period = 0; sample_start adds the count-0 sample) or on a time
    grid; returns a dict with S, M, time, weight (and, by pass, the total rate)
    per sample, final occ, engine, the class array, final time and the pending
    event (site, dt)."""
    n = rows * cols
    o = np.ascontiguousarray(occ, dtype=np.int8).copy()
    eng = np.ascontiguousarray(engine, dtype=np.uint64).copy()
    m = capacity if capacity is not None else n_passes // sample_period + (1 if sample_start else 0)
    S = np.zeros(max(m, 1), dtype=np.int64)
    M = np.zeros(max(m, 1), dtype=np.int64)
    t = np.zeros(max(m, 1), dtype=np.float64)
    w = np.zeros(max(m, 1), dtype=np.float64)
    rate = np.zeros(max(m, 1), dtype=np.float64)
    arr = np.zeros(n, dtype=np.int64)
    sc = np.zeros(2, dtype=np.float64)
    pend = C.c_int64(0)
    got = lib().orc_nfold_run(rows, cols, J, T, mu, n_passes, sample_period, time_period, m,
                              _p(o, C.c_int8), _p(eng, C.c_uint64), _p(S, C.c_int64),
                              _p(M, C.c_int64), _p(t, C.c_double), _p(w, C.c_double),
                              _p(arr, C.c_int64), _p(sc, C.c_double), C.byref(pend),
                              1 if sample_start else 0, _p(rate, C.c_double))
    m = min(m, got)
    return dict(S=S[:m], M=M[:m], time=t[:m], weight=w[:m], rate=rate[:m], occ=o, engine=eng,
                arr=arr,
                n_total=got,
                final_time=float(sc[0]), pending_dt=float(sc[1]), pending_site=pend.value)
