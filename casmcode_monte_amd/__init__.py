"""casmcode_monte_amd -- MI355X-native occupation-Metropolis hot path of libcasm-monte.

The compute path is the HIP library lib/libcasm_monte_mi355x.so, reached
through the C ABI of include/casm_mi355x/mcgpu.h.  This module is a thin
ctypes host binding of that ABI (used by the tests and bench.py; the C++
drop-in classes under include/casm/monte/ bind the same ABI).  There is no CPU
fallback: loading fails loudly when the library or a gfx950 device is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcasm_monte_mi355x.so")
INCLUDE_DIR = os.path.join(os.path.dirname(HERE), "include")

MODE_EXACT = 0
MODE_CHECKERBOARD = 1
MODE_CANONICAL = 2
MODE_NFOLD = 3
MODE_OCCUPANT = 4
MODE_KMC = 5
LATTICE_SQUARE = 2
LATTICE_SC = 3
LATTICE_FCC = 4
OBS_FORMATION_ENERGY = 0
OBS_PARAM_COMPOSITION = 1
OBS_POTENTIAL_ENERGY = 2
OPT_FORCE_STREAMING = 1
OPT_EXACT_SCALAR_WALK = 2
OPT_CHECKER_QUAD_PATH = 4
OPT_NFOLD_SAMPLE_START = 8

# Boltzmann constant, eV/K, as compiled into CASMcode_global (used by the
# reference's beta = 1/(KB*T), basic_occupation_metropolis.hh:361).
KB = 8.6173303e-05

EXPORTED = [
    "mcg_create", "mcg_destroy", "mcg_last_error", "mcg_set_occupation", "mcg_get_occupation",
    "mcg_rng_dump", "mcg_rng_load", "mcg_run", "mcg_set_active", "mcg_get_counters",
    "mcg_n_samples", "mcg_get_samples", "mcg_clear_samples", "mcg_stats_compute",
    "mcg_stats_weighted", "mcg_stats_compute_noalloc",
    "mcg_equilibration", "mcg_last_run_time", "mcg_assign_chain", "mcg_autotune",
    "mcg_get_schedule", "mcg_set_schedule", "mcg_reserve_samples", "mcg_get_occ_location",
    "mcg_set_occ_location", "mcg_get_nfold_state", "mcg_run_steps", "mcg_get_position",
    "mcg_check_batch", "mcg_set_site_asym", "mcg_get_atoms", "mcg_kmc_set_spin_limit",
    "mcg_kmc_advance", "mcg_fcc_class_schedule", "mcg_kmc_collect_hops", "mcg_kmc_get_hops",
]


class ChainDesc(C.Structure):
    _fields_ = [("shape", C.c_int64 * 3), ("temperature", C.c_double),
                ("exchange_potential", C.c_double), ("seed", C.c_uint64)]


class OccupantDesc(C.Structure):
    _fields_ = [("n_asym", C.c_int64), ("occ_size", C.POINTER(C.c_int64)),
                ("species", C.POINTER(C.c_int64)), ("n_species", C.c_int64),
                ("canonical", C.c_int64), ("n_swaps", C.c_int64),
                ("swaps", C.POINTER(C.c_int64)), ("exchange_potential", C.POINTER(C.c_double)),
                ("sites_per_unitcell", C.c_int64)]


class Desc(C.Structure):
    _fields_ = [("mode", C.c_int64), ("lattice", C.c_int64), ("use_nlist", C.c_int64),
                ("J1", C.c_double), ("J2", C.c_double), ("kb", C.c_double),
                ("sample_period", C.c_int64), ("n_chains", C.c_int64),
                ("chains", C.POINTER(ChainDesc)), ("options", C.c_int64),
                ("eci", C.POINTER(C.c_double)), ("sample_time_period", C.c_double),
                ("occupant", C.POINTER(OccupantDesc))]


SAMPLE_DTYPE = np.dtype([("pass", np.int64), ("S1", np.int64), ("S2", np.int64),
                         ("M", np.int64), ("e_fold", np.float64), ("e_run", np.float64),
                         ("S3", np.int64), ("T3", np.int64), ("time", np.float64),
                         ("weight", np.float64)])


class Schedule(C.Structure):
    _fields_ = [("lds_block", C.c_int64), ("units_per_thread", C.c_int64),
                ("graph_passes", C.c_int64), ("trial_ms", C.c_double),
                ("band_lines", C.c_int64), ("line_tiles", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("mean", C.c_double), ("calculated_precision", C.c_double),
                ("variance", C.c_double), ("autocorr_factor", C.c_double), ("n", C.c_int64)]


class Check(C.Structure):
    _fields_ = [("is_equilibrated", C.c_int64), ("n_equilibration", C.c_int64),
                ("stats", Stats)]


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.mcg_create.argtypes = [P(Desc), P(C.c_void_p)]
        L.mcg_destroy.argtypes = [C.c_void_p]
        L.mcg_destroy.restype = None
        L.mcg_last_error.restype = C.c_char_p
        L.mcg_set_occupation.argtypes = [C.c_void_p, C.c_int64, P(C.c_int8), C.c_int64]
        L.mcg_get_occupation.argtypes = [C.c_void_p, C.c_int64, P(C.c_int8), C.c_int64]
        L.mcg_rng_dump.argtypes = [C.c_void_p, C.c_int64, P(C.c_uint64), C.c_int64]
        L.mcg_rng_load.argtypes = [C.c_void_p, C.c_int64, P(C.c_uint64), C.c_int64]
        L.mcg_run.argtypes = [C.c_void_p, C.c_int64]
        L.mcg_set_active.argtypes = [C.c_void_p, P(C.c_uint8), C.c_int64]
        L.mcg_get_counters.argtypes = [C.c_void_p, C.c_int64, P(C.c_int64), P(C.c_int64),
                                       P(C.c_int64)]
        L.mcg_n_samples.argtypes = [C.c_void_p, C.c_int64, P(C.c_int64)]
        L.mcg_get_nfold_state.argtypes = [C.c_void_p, C.c_int64, P(C.c_double), P(C.c_int64),
                                          P(C.c_double)]
        L.mcg_get_samples.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        L.mcg_clear_samples.argtypes = [C.c_void_p]
        L.mcg_stats_compute.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                                        P(Stats)]
        L.mcg_equilibration.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_double,
                                        P(C.c_int64), P(C.c_int64)]
        L.mcg_stats_weighted.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                                         C.c_int64, C.c_int64, P(Stats)]
        L.mcg_stats_compute_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64,
                                                C.c_double, P(Stats)]
        L.mcg_run_steps.argtypes = [C.c_void_p, C.c_int64]
        L.mcg_check_batch.argtypes = [C.c_void_p, C.c_int64, P(C.c_int64), C.c_int64, C.c_double,
                                      C.c_double, P(Check)]
        L.mcg_get_position.argtypes = [C.c_void_p, C.c_int64, P(C.c_int64), P(C.c_int64),
                                       P(C.c_double)]
        L.mcg_last_run_time.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
        L.mcg_assign_chain.argtypes = [C.c_void_p, C.c_int64, P(ChainDesc)]
        L.mcg_autotune.argtypes = [C.c_void_p, C.c_int64, P(Schedule)]
        L.mcg_reserve_samples.argtypes = [C.c_void_p, C.c_int64]
        L.mcg_get_schedule.argtypes = [C.c_void_p, P(Schedule)]
        L.mcg_set_schedule.argtypes = [C.c_void_p, P(Schedule)]
        L.mcg_get_occ_location.argtypes = [C.c_void_p, C.c_int64, P(C.c_int64), C.c_int64]
        L.mcg_set_occ_location.argtypes = [C.c_void_p, C.c_int64, P(C.c_int64), C.c_int64]
        L.mcg_set_site_asym.argtypes = [C.c_void_p, C.c_int64, P(C.c_int8), C.c_int64]
        L.mcg_get_atoms.argtypes = [C.c_void_p, C.c_int64, P(C.c_int64), C.c_int64]
        L.mcg_fcc_class_schedule.argtypes = [C.c_int64, C.c_int64, C.c_int64, P(C.c_int64),
                                             P(C.c_int64), P(C.c_int64)]
        L.mcg_kmc_collect_hops.argtypes = [C.c_void_p, C.c_int64, C.c_int64]
        L.mcg_kmc_get_hops.argtypes = [C.c_void_p, C.c_int64, C.c_int64, P(C.c_int64),
                                       P(C.c_int64), P(C.c_double)]
        L.mcg_kmc_set_spin_limit.argtypes = [C.c_void_p, C.c_int64]
        L.mcg_kmc_advance.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_int64]
        _lib = L
    return _lib


class McgError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise McgError(f"mcg error {rc}: {lib().mcg_last_error().decode()}")


def fcc_class_schedule(shape):
    """The class schedule an FCC chain of shape (L0, L1, L2) conventional cells
    sweeps each pass: a list of steps, each a list of colour classes
    (col = cx + 6 cy + 36 cz) that update together (a host computation)."""
    n = C.c_int64()
    start = (C.c_int64 * 129)()
    cls = (C.c_int64 * 128)()
    _check(lib().mcg_fcc_class_schedule(int(shape[0]), int(shape[1]), int(shape[2]),
                                        C.byref(n), start, cls))
    return [[cls[k] for k in range(start[s], start[s + 1])] for s in range(n.value)]


class Run:
    """One handle of the C ABI: a batch of chains advanced together on the GPU."""

    def __init__(self, chains, mode=MODE_EXACT, lattice=LATTICE_SQUARE, J1=0.1, J2=0.0,
                 kb=KB, use_nlist=True, sample_period=1, options=0, eci=None,
                 sample_time_period=0.0, occupant=None):
        """chains: sequence of dicts with shape, temperature, exchange_potential, seed.
        FCC (lattice=LATTICE_FCC, mode CHECKERBOARD): shape in conventional cells and
        eci = (V0, V1, V2, V3, V4, Vt).  OCCUPANT: occupant = dict(asym_species=
        [[species of occupant 0, occupant 1], ...] per unit (one entry for a
        unit with one occupant), n_species, canonical (bool), swaps = [(cand_a,
        cand_b), ...], exchange_potential = per chain a list of n_species values
        (optional), sites_per_unitcell (optional))."""
        L = lib()
        self.n_chains = len(chains)
        arr = (ChainDesc * self.n_chains)()
        self.shapes = []
        for i, ch in enumerate(chains):
            shp = list(ch["shape"]) + [1] * (3 - len(ch["shape"]))
            arr[i].shape[:] = shp
            arr[i].temperature = ch["temperature"]
            arr[i].exchange_potential = ch.get("exchange_potential", 0.0)
            arr[i].seed = ch.get("seed", 0)
            self.shapes.append(tuple(shp))
        self.chains = [dict(c) for c in chains]
        self.eci = None
        eci_arr = None
        if lattice == LATTICE_FCC:
            if eci is None or len(eci) != 6:
                raise ValueError("the FCC cluster expansion needs 6 ECIs")
            self.eci = tuple(float(v) for v in eci)
            eci_arr = (C.c_double * 6)(*self.eci)
        occ_ptr = None
        self._occ_keep = None
        self.occupant = None
        if mode == MODE_OCCUPANT:
            if occupant is None:
                raise ValueError("OCCUPANT mode needs the occupant system")
            spec = occupant["asym_species"]
            na = len(spec)
            osz = (C.c_int64 * na)(*[len(x) for x in spec])
            sp = (C.c_int64 * (2 * na))(*[v for x in spec for v in (list(x) + [0])[:2]])
            sw = [v for pr in occupant["swaps"] for v in pr]
            swa = (C.c_int64 * len(sw))(*sw)
            ns = int(occupant["n_species"])
            mu = occupant.get("exchange_potential")
            mua = None
            if mu is not None:
                flat = [float(v) for row in mu for v in row]
                if len(flat) != self.n_chains * ns:
                    raise ValueError("exchange_potential: n_chains x n_species values")
                mua = (C.c_double * len(flat))(*flat)
            od = OccupantDesc(na, osz, sp, ns, int(bool(occupant.get("canonical", False))),
                              len(sw) // 2, swa, mua, int(occupant.get("sites_per_unitcell", 1)))
            self._occ_keep = (osz, sp, swa, mua, od)
            occ_ptr = C.pointer(od)
            self.occupant = dict(occupant)
        d = Desc(mode, lattice, int(use_nlist), J1, J2, kb, sample_period, self.n_chains, arr,
                 options, eci_arr, sample_time_period, occ_ptr)
        self._arr, self._eci_arr = arr, eci_arr
        h = C.c_void_p()
        _check(L.mcg_create(C.byref(d), C.byref(h)))
        self.h = h
        self.mode, self.lattice, self.J1, self.J2 = mode, lattice, J1, J2
        self.use_nlist, self.sample_period, self.kb = use_nlist, sample_period, kb

    def close(self):
        if getattr(self, "h", None):
            lib().mcg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def n_sites(self, c):
        s = self.shapes[c]
        return (4 if self.lattice == LATTICE_FCC else 1) * s[0] * s[1] * s[2]

    def set_occupation(self, c, occ):
        o = np.ascontiguousarray(occ, dtype=np.int8)
        _check(lib().mcg_set_occupation(self.h, c, o.ctypes.data_as(C.POINTER(C.c_int8)), o.size))

    def get_occupation(self, c):
        o = np.zeros(self.n_sites(c), dtype=np.int8)
        _check(lib().mcg_get_occupation(self.h, c, o.ctypes.data_as(C.POINTER(C.c_int8)), o.size))
        return o

    def set_site_asym(self, c, asym):
        """OCCUPANT: the asymmetric unit of every site (resets the occupation to
        occupant 0 and the lists)."""
        a = np.ascontiguousarray(asym, dtype=np.int8)
        _check(lib().mcg_set_site_asym(self.h, c, a.ctypes.data_as(C.POINTER(C.c_int8)), a.size))
        self._units = getattr(self, "_units", {})
        self._units[c] = a.copy()

    def atoms(self, c, n_atoms):
        """KMC: (n_atoms, 4) array of site, displacement along rows and columns,
        jumps, in atom order."""
        v = np.zeros(4 * n_atoms, dtype=np.int64)
        _check(lib().mcg_get_atoms(self.h, c, v.ctypes.data_as(C.POINTER(C.c_int64)), n_atoms))
        return v.reshape(n_atoms, 4)

    def kmc_collect_hops(self, jumps_per_position_sample, max_n_position_samples=100):
        """KMC: record hop correlations on the device (0 turns recording off)."""
        self._max_hops = int(max_n_position_samples)
        _check(lib().mcg_kmc_collect_hops(self.h, int(jumps_per_position_sample),
                                          int(max_n_position_samples)))

    def kmc_hops(self, c, n_atoms):
        """KMC: (counts[n_atoms], records[max, n_atoms, 4] = dx, dy, step, pass,
        times[max, n_atoms]) of chain c."""
        m = self._max_hops
        counts = np.zeros(n_atoms, dtype=np.int64)
        rec = np.zeros(m * n_atoms * 4, dtype=np.int64)
        t = np.zeros(m * n_atoms, dtype=np.float64)
        _check(lib().mcg_kmc_get_hops(self.h, c, n_atoms,
                                      counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                      rec.ctypes.data_as(C.POINTER(C.c_int64)),
                                      t.ctypes.data_as(C.POINTER(C.c_double))))
        return counts, rec.reshape(m, n_atoms, 4), t.reshape(m, n_atoms)

    def kmc_advance(self, max_events, stop_time=float("inf"), select_at_end=False):
        """KMC: apply the pending event and events up to max_events; an event
        reaching stop_time stays pending; select_at_end leaves one more selected
        (mcg_kmc_advance)."""
        _check(lib().mcg_kmc_advance(self.h, int(max_events), float(stop_time),
                                     int(bool(select_at_end))))

    def kmc_set_spin_limit(self, attempts):
        _check(lib().mcg_kmc_set_spin_limit(self.h, int(attempts)))

    def n_mutating(self, c):
        """OCCUPANT: sites whose unit allows two occupants."""
        units = getattr(self, "_units", {}).get(c)
        if units is None:
            units = np.zeros(self.n_sites(c), dtype=np.int8)
        spec = self.occupant["asym_species"]
        return int(sum(1 for u in units if len(spec[u]) > 1))

    def occ_location(self, c):
        """CANONICAL: the candidate lists (species-0 sites in list order, then species 1).
        OCCUPANT: the candidate lists in candidate order, as site indices."""
        n = self.n_mutating(c) if self.mode == MODE_OCCUPANT else self.n_sites(c)
        v = np.zeros(n, dtype=np.int64)
        _check(lib().mcg_get_occ_location(self.h, c, v.ctypes.data_as(C.POINTER(C.c_int64)),
                                          v.size))
        return v

    def set_occ_location(self, c, loc):
        v = np.ascontiguousarray(loc, dtype=np.int64)
        _check(lib().mcg_set_occ_location(self.h, c, v.ctypes.data_as(C.POINTER(C.c_int64)),
                                          v.size))

    def rng_dump(self, c):
        n = 4 if self.mode == MODE_CHECKERBOARD else 313
        w = np.zeros(n, dtype=np.uint64)
        _check(lib().mcg_rng_dump(self.h, c, w.ctypes.data_as(C.POINTER(C.c_uint64)), n))
        return w

    def rng_load(self, c, words):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        _check(lib().mcg_rng_load(self.h, c, w.ctypes.data_as(C.POINTER(C.c_uint64)), w.size))

    def run(self, n_passes):
        """Run n_passes passes of every active chain.  NFOLD sampled by pass:
        run(0) selects the next event of each chain that has none pending and
        records the current pass (its state, the event's time increment as
        weight, the total rate in e_run); the next run applies that event first."""
        _check(lib().mcg_run(self.h, int(n_passes)))

    def run_steps(self, n_steps):
        """EXACT, OCCUPANT, KMC: exactly n_steps more attempts (KMC: events) per
        active chain (may end inside a pass)."""
        _check(lib().mcg_run_steps(self.h, n_steps))

    def position(self, c):
        """(pass count, steps into the current pass, FP64 running energy per unit cell)."""
        p, st, e = C.c_int64(), C.c_int64(), C.c_double()
        _check(lib().mcg_get_position(self.h, c, C.byref(p), C.byref(st), C.byref(e)))
        return p.value, st.value, e.value

    def assign(self, c, chain):
        """Re-initialise slot c for a new run (shape within the slot's capacity)."""
        shp = list(chain["shape"]) + [1] * (3 - len(chain["shape"]))
        d = ChainDesc()
        d.shape[:] = shp
        d.temperature = chain["temperature"]
        d.exchange_potential = chain.get("exchange_potential", 0.0)
        d.seed = chain.get("seed", 0)
        _check(lib().mcg_assign_chain(self.h, c, C.byref(d)))
        getattr(self, "_units", {}).pop(c, None)
        self.shapes[c] = tuple(shp)
        self.chains[c] = dict(chain)

    def set_active(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        _check(lib().mcg_set_active(self.h, m.ctypes.data_as(C.POINTER(C.c_uint8)), m.size))

    def counters(self, c):
        a, b, d = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().mcg_get_counters(self.h, c, C.byref(a), C.byref(b), C.byref(d)))
        return a.value, b.value, d.value

    def nfold_state(self, c):
        """NFOLD: (simulated time, pending event site or -1, its time increment)."""
        t, site, dt = C.c_double(), C.c_int64(), C.c_double()
        _check(lib().mcg_get_nfold_state(self.h, c, C.byref(t), C.byref(site), C.byref(dt)))
        return t.value, site.value, dt.value

    def n_samples(self, c):
        n = C.c_int64()
        _check(lib().mcg_n_samples(self.h, c, C.byref(n)))
        return n.value

    def samples(self, c, first=0, count=None):
        n = self.n_samples(c)
        if count is None:
            count = n - first
        out = np.zeros(count, dtype=SAMPLE_DTYPE)
        _check(lib().mcg_get_samples(self.h, c, first, count, out.ctypes.data_as(C.c_void_p)))
        return out

    def clear_samples(self):
        _check(lib().mcg_clear_samples(self.h))

    def stats(self, c, observable, first=0, confidence=0.95):
        s = Stats()
        _check(lib().mcg_stats_compute(self.h, c, observable, first, confidence, C.byref(s)))
        return dict(mean=s.mean, calculated_precision=s.calculated_precision,
                    variance=s.variance, autocorr_factor=s.autocorr_factor, n=s.n)

    def stats_noalloc(self, c, observable, first=0, confidence=0.95):
        """stats() computed in the library's fixed device scratch
        (mcg_stats_compute_noalloc: no device allocation; at most
        MCG_STATS_NOALLOC_MAX - 8 observations)."""
        s = Stats()
        _check(lib().mcg_stats_compute_noalloc(self.h, c, observable, first, confidence,
                                               C.byref(s)))
        return dict(mean=s.mean, calculated_precision=s.calculated_precision,
                    variance=s.variance, autocorr_factor=s.autocorr_factor, n=s.n)

    def stats_weighted(self, c, observable, first=0, confidence=0.95, method=1,
                       n_resamples=10000):
        """Statistics of weighted (NFOLD) records: BasicStatisticsCalculator's
        weighted branch (BasicStatistics.cc:144-188)."""
        s = Stats()
        _check(lib().mcg_stats_weighted(self.h, c, observable, first, confidence, method,
                                        n_resamples, C.byref(s)))
        return dict(mean=s.mean, calculated_precision=s.calculated_precision,
                    variance=s.variance, autocorr_factor=s.autocorr_factor, n=s.n)

    def check_batch(self, chains, observable, prec, confidence=0.95):
        """Equilibration at precision prec, then the statistics after it, for every
        listed chain in one launch (mcg_check_batch); a list of dicts."""
        ch = np.ascontiguousarray(chains, dtype=np.int64)
        out = (Check * max(len(ch), 1))()
        _check(lib().mcg_check_batch(self.h, len(ch), ch.ctypes.data_as(C.POINTER(C.c_int64)),
                                     observable, prec, confidence, out))
        res = []
        for k in range(len(ch)):
            s = out[k].stats
            res.append(dict(is_equilibrated=bool(out[k].is_equilibrated),
                            n_equilibration=out[k].n_equilibration, mean=s.mean,
                            calculated_precision=s.calculated_precision, variance=s.variance,
                            autocorr_factor=s.autocorr_factor, n=s.n))
        return res

    def equilibration(self, c, observable, prec):
        a, b = C.c_int64(), C.c_int64()
        _check(lib().mcg_equilibration(self.h, c, observable, prec, C.byref(a), C.byref(b)))
        return bool(a.value), b.value

    def reserve_samples(self, n_records):
        """Grow every chain's device sample series to n_records records."""
        _check(lib().mcg_reserve_samples(self.h, int(n_records)))

    def autotune(self, trial_passes=64):
        """Time the candidate launch schedules on this run's chains (state is
        restored afterwards) and keep the fastest; returns it as a dict."""
        s = Schedule()
        _check(lib().mcg_autotune(self.h, trial_passes, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Schedule._fields_}

    def schedule(self):
        s = Schedule()
        _check(lib().mcg_get_schedule(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Schedule._fields_}

    def set_schedule(self, lds_block=None, units_per_thread=None, graph_passes=None,
                     band_lines=None, line_tiles=None):
        cur = self.schedule()
        s = Schedule(lds_block if lds_block is not None else cur["lds_block"],
                     units_per_thread if units_per_thread is not None
                     else cur["units_per_thread"],
                     graph_passes if graph_passes is not None else cur["graph_passes"], 0.0,
                     band_lines if band_lines is not None else cur["band_lines"],
                     line_tiles if line_tiles is not None else cur["line_tiles"])
        _check(lib().mcg_set_schedule(self.h, C.byref(s)))

    def last_run_time(self):
        ms, att = C.c_double(), C.c_double()
        _check(lib().mcg_last_run_time(self.h, C.byref(ms), C.byref(att)))
        return ms.value, att.value

    # -- the reference's FP64 sampling functions, recomputed from exact integers --
    def observables(self, c, rec=None):
        """formation_energy, param_composition, potential_energy per sample, with the
        reference's rounding (model.hh:266-295, 412-422; basic_semigrand_canonical.hh:165-174)."""
        if rec is None:
            rec = self.samples(c)
        N = np.float64(self.n_sites(c))
        mu = np.float64(self.chains[c].get("exchange_potential", 0.0))
        if self.lattice == LATTICE_FCC:
            # E_f = N V0 + V1 N1 + V2/6 P1 + V3/3 P2 + V4/12 P3 + Vt/8 T (device order)
            v = [np.float64(x) for x in self.eci]
            coef = [v[0], v[1], v[2] / 6.0, v[3] / 3.0, v[4] / 12.0, v[5] / 8.0]
            Ef = coef[0] * N
            for k, f in zip(coef[1:], ("M", "S1", "S2", "S3", "T3")):
                Ef = Ef + k * rec[f].astype(np.float64)
            N1 = rec["M"].astype(np.float64)
            return dict(formation_energy=Ef / N, param_composition=N1 / N,
                        potential_energy=(Ef - mu * N1) / N)
        if self.mode in (MODE_EXACT, MODE_CANONICAL) and not self.use_nlist:
            Ef = rec["e_fold"].astype(np.float64)
        else:
            Ef = rec["S1"].astype(np.float64) * np.float64(-self.J1)
            if self.J2 != 0.0:
                Ef = Ef + rec["S2"].astype(np.float64) * np.float64(-self.J2)
        Nx = (np.int64(self.n_sites(c)) + rec["M"]).astype(np.float64) / np.float64(2.0)
        mu = np.float64(self.chains[c].get("exchange_potential", 0.0))
        if self.mode == MODE_CANONICAL:
            mu = np.float64(0.0)  # the canonical potential is the formation energy
        return dict(formation_energy=Ef / N, param_composition=Nx / N,
                    potential_energy=(Ef - mu * Nx) / N)
