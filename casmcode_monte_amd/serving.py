"""Continuous-batching server of SGC Metropolis runs on one MI355X.

A request is one run of the 2D square-lattice Ising model (edge L) or of the
binary FCC cluster expansion (L^3 conventional cells), at a temperature,
exchange potential and seed, that stops when the precision of <x> (the
parametric composition) reaches the requested value or after max_passes.
Stopping follows libcasm-monte's completion check with one requested
precision (checks/CompletionCheck.hh:272-376): at every check the series is
tested for equilibration (EquilibrationCheck.cc:50-117, precision = the
requested absolute precision) and, once equilibrated, the tail after the
equilibration samples must reach calculated_precision < requested
(ConvergenceCheck.hh:74-119, BasicStatistics.cc:114-131).  Both reductions run
on the device over the device-resident sample series.

Slots: requests are routed by model and lattice size to pools (one library
handle and one host thread each): LDS-resident classes up to 128 KB Ising
lattices and, for FCC, every class (the bit-line kernel holds a 48^3-cell
lattice in 92 KB of LDS), and an HBM class above 362 for Ising.  A pool runs `chunk` passes of
all its active slots per round (one launch), then checks the slots whose own
pass count reached a multiple
of `check_period`, retires finished requests and admits waiting ones into the
freed slots (mcg_assign_chain) -- requests that arrive while others run join
the next round.  Admission also respects a global HBM footprint budget
(lattice bytes + sample-series bytes of every admitted request).

Every slot advances in whole chunks of its own passes and the CHECKERBOARD
random stream depends only on (seed, pass, colour, site), so a request's
result does not depend on what else is in the batch: alone and in the mix it
is bit-identical.
"""
import dataclasses
import threading
import time
import zlib

import numpy as np

from . import (KB, LATTICE_FCC, LATTICE_SQUARE, MODE_CHECKERBOARD, OBS_PARAM_COMPOSITION,
               OPT_FORCE_STREAMING, SAMPLE_DTYPE, Run)
from .fcc import DEFAULT_ECI


@dataclasses.dataclass
class Request:
    rid: int
    edge: int
    temperature: float
    exchange_potential: float
    seed: int
    arrival: float = 0.0       # seconds after the start of the trace
    max_passes: int = 20000
    precision: float = 1e-3    # requested absolute precision of <x>
    model: str = "ising"       # "ising" (2D, edge L) or "fcc" (L^3 conventional cells)

    def sites(self):
        return 4 * self.edge ** 3 if self.model == "fcc" else self.edge * self.edge

    def shape(self):
        return (self.edge,) * 3 if self.model == "fcc" else (self.edge, self.edge)

    def footprint(self, sample_bytes=SAMPLE_DTYPE.itemsize):
        return self.sites() + self.max_passes * sample_bytes


@dataclasses.dataclass
class Result:
    rid: int
    passes: int
    n_samples: int
    converged: bool
    n_equilibration: int
    mean_x: float
    precision_x: float
    lattice_crc: int
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self):
        return self.finished - self.arrival


# (max edge, slots, LDS-resident) per size class
DEFAULT_CLASSES = ((64, 96, True), (128, 64, True), (256, 48, True), (362, 32, True),
                   (512, 16, False))
# FCC: edges in conventional cells (4 L^3 sites, bit lines in LDS up to L = 48)
DEFAULT_FCC_CLASSES = ((12, 64, True), (20, 48, True), (32, 32, True), (48, 12, True))


class _Pool:
    def __init__(self, model, max_edge, n_slots, lds, J1, J2, eci, chunk, check_period):
        self.model = model
        self.max_edge = max_edge
        opts = 0 if lds else OPT_FORCE_STREAMING
        if model == "fcc":
            chains = [dict(shape=(max_edge,) * 3, temperature=1000.0, exchange_potential=0.0,
                           seed=i) for i in range(n_slots)]
            self.run = Run(chains, mode=MODE_CHECKERBOARD, lattice=LATTICE_FCC, eci=eci,
                           options=opts)
        else:
            chains = [dict(shape=(max_edge, max_edge), temperature=1000.0,
                           exchange_potential=0.0, seed=i) for i in range(n_slots)]
            self.run = Run(chains, mode=MODE_CHECKERBOARD, lattice=LATTICE_SQUARE, J1=J1,
                           J2=J2, options=opts)
        self.active = np.zeros(n_slots, dtype=np.uint8)
        self.run.set_active(self.active)
        self.slot_req = [None] * n_slots
        self.passes = np.zeros(n_slots, dtype=np.int64)
        self.admitted = np.zeros(n_slots)
        self.chunk, self.check_period = chunk, check_period
        self.queue = []   # requests of this class, in arrival order
        self.attempts = 0.0


class Server:
    """Continuous-batching Metropolis server (CHECKERBOARD mode: 2D square-lattice
    Ising and binary FCC cluster expansion pools)."""

    def __init__(self, classes=DEFAULT_CLASSES, J1=0.1, J2=0.0, chunk=100, check_period=500,
                 hbm_budget=8 << 30, fcc_classes=(), eci=DEFAULT_ECI):
        if check_period % chunk:
            raise ValueError("check_period must be a multiple of chunk")
        self.pools = [_Pool("ising", e, n, lds, J1, J2, None, chunk, check_period)
                      for e, n, lds in classes]
        self.pools += [_Pool("fcc", e, n, lds, J1, J2, eci, chunk, check_period)
                       for e, n, lds in fcc_classes]
        self.hbm_budget = hbm_budget
        self.in_use = 0
        self.max_in_use = 0
        self.lock = threading.Lock()
        self.results = {}
        self.errors = []

    def pool_for(self, req):
        if req.model == "ising" and req.edge < 2:
            raise ValueError("Ising edge lengths must be >= 2")
        if req.model == "fcc" and req.edge < 3:
            raise ValueError("FCC edges must be >= 3 conventional cells")
        for p in self.pools:
            if p.model == req.model and req.edge <= p.max_edge:
                return p
        raise ValueError(f"no {req.model} slot class holds edge {req.edge}")

    # --- one pool's loop ----------------------------------------------------------
    def _admit(self, p, now):
        changed = False
        while p.queue and p.queue[0].arrival <= now:
            free = np.flatnonzero(p.active == 0)
            if free.size == 0:
                break
            req = p.queue[0]
            with self.lock:
                if self.in_use + req.footprint() > self.hbm_budget:
                    break
                self.in_use += req.footprint()
                self.max_in_use = max(self.max_in_use, self.in_use)
            p.queue.pop(0)
            s = int(free[0])
            p.run.assign(s, dict(shape=req.shape(), temperature=req.temperature,
                                 exchange_potential=req.exchange_potential, seed=req.seed))
            p.slot_req[s] = req
            p.passes[s] = 0
            p.admitted[s] = now
            p.active[s] = 1
            changed = True
        if changed:
            p.run.set_active(p.active)

    def _check_due(self, p, due, t0):
        """Completion checks of the due slots, batched per requested precision
        (mcg_check_batch: one launch for all of them); returns {slot: Result} for
        the requests that are done."""
        done = {}
        by_prec = {}
        for s in due:
            by_prec.setdefault(p.slot_req[s].precision, []).append(int(s))
        for prec, slots in by_prec.items():
            checks = p.run.check_batch(slots, OBS_PARAM_COMPOSITION, prec)
            for s, ck in zip(slots, checks):
                res = self._result(p, s, t0, ck)
                if res is not None:
                    done[s] = res
        return done

    def _result(self, p, s, t0, ck):
        """The Result of slot s from its completion check, or None if it goes on."""
        req = p.slot_req[s]
        n = p.run.n_samples(s)
        is_eq, n_eq = ck["is_equilibrated"], ck["n_equilibration"]
        converged, mean, prec = False, float("nan"), float("inf")
        if is_eq and n_eq < n:
            mean, prec = ck["mean"], ck["calculated_precision"]
            converged = prec < req.precision
        if not converged and p.passes[s] < req.max_passes:
            return None
        if not is_eq or n_eq >= n:  # report the plain mean at the cutoff
            x = p.run.observables(s)["param_composition"]
            mean = float(np.mean(x))
        occ = p.run.get_occupation(s)
        return Result(req.rid, int(p.passes[s]), n, converged, int(n_eq), mean, prec,
                      zlib.crc32(occ.tobytes()), req.arrival, float(p.admitted[s]),
                      time.perf_counter() - t0)

    def _loop(self, p, t0):
        try:
            while p.queue or p.active.any():
                now = time.perf_counter() - t0
                self._admit(p, now)
                if not p.active.any():
                    wait = p.queue[0].arrival - now if p.queue else 0.0
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                    continue
                p.run.run(p.chunk)
                ms, att = p.run.last_run_time()
                p.attempts += att
                p.passes[p.active == 1] += p.chunk
                retired = False
                due = [s for s in np.flatnonzero(p.active)
                       if p.passes[s] % p.check_period == 0
                       or p.passes[s] >= p.slot_req[s].max_passes]
                for s, res in self._check_due(p, due, t0).items():
                    req = p.slot_req[s]
                    self.results[req.rid] = res
                    with self.lock:
                        self.in_use -= req.footprint()
                    p.active[s] = 0
                    p.slot_req[s] = None
                    retired = True
                if retired:
                    p.run.set_active(p.active)
        except Exception as e:  # surfaced by serve()
            self.errors.append(e)

    def serve(self, requests):
        """Replay a request trace (arrival times in seconds from now); returns the
        results by request id and the run statistics."""
        for p in self.pools:
            p.queue = []
            p.attempts = 0.0
        for r in sorted(requests, key=lambda r: (r.arrival, r.rid)):
            self.pool_for(r).queue.append(r)
        t0 = time.perf_counter()
        threads = [threading.Thread(target=self._loop, args=(p, t0)) for p in self.pools]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        if self.errors:
            raise self.errors[0]
        lat = np.array([r.latency for r in self.results.values()])
        attempts = sum(p.attempts for p in self.pools)
        return self.results, dict(
            wall_s=wall, attempts=attempts, attempts_per_s=attempts / wall,
            p50_latency_s=float(np.percentile(lat, 50)) if lat.size else None,
            p99_latency_s=float(np.percentile(lat, 99)) if lat.size else None,
            n_requests=len(self.results), max_hbm_in_use=self.max_in_use)

    def close(self):
        for p in self.pools:
            p.run.close()


def poisson_trace(n, rate=50.0, seed=2024, edge_range=(16, 512), T_range=(300.0, 3000.0),
                  mu_range=(-1.0, 1.0), max_passes=20000, precision=1e-3, fcc_fraction=0.0,
                  fcc_edge_range=(8, 48)):
    """Request trace: Poisson arrivals at `rate` per second, T and mu uniform.
    Ising edges are log-uniform in edge_range rounded to the nearest integer
    (odd edges run with ring colourings); with fcc_fraction > 0 each request is
    an FCC run with that
    probability (drawn after the Ising trace, so fcc_fraction = 0 reproduces it),
    its edge log-uniform in fcc_edge_range rounded to the nearest integer."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    lo, hi = np.log(edge_range[0]), np.log(edge_range[1])
    edges = np.exp(rng.uniform(lo, hi, size=n))
    edges = np.clip(np.round(edges), edge_range[0], edge_range[1]).astype(int)
    T = rng.uniform(*T_range, size=n)
    mu = rng.uniform(*mu_range, size=n)
    models = ["ising"] * n
    if fcc_fraction > 0.0:
        is_fcc = rng.random(n) < fcc_fraction
        flo, fhi = np.log(fcc_edge_range[0]), np.log(fcc_edge_range[1])
        fedges = np.clip(np.round(np.exp(rng.uniform(flo, fhi, size=n))), *fcc_edge_range)
        for i in np.flatnonzero(is_fcc):
            models[i] = "fcc"
            edges[i] = int(fedges[i])
    return [Request(i, int(edges[i]), float(T[i]), float(mu[i]), 0x5E00000 + i, float(t[i]),
                    max_passes, precision, models[i]) for i in range(n)]
