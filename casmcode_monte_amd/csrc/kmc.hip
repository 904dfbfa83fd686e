// kmc.hip -- KMC mode: kinetic Monte Carlo with atom tracking for a square
// lattice gas, bit for bit the loop of kinetic_monte_carlo.hh:136-229 under a
// rejection event selector, on one wavefront per chain; a second wavefront of
// the workgroup twists the engine's next generation meanwhile (mt_producer,
// walk.hh).
//
// Model: occupant 0 = vacancy, 1 = atom, E = -J sum over nearest-neighbour
// bonds of n_i n_j.  Selector: attempts follow one another at 1/R, R = 4 *
// (number of atoms); an attempt chooses an atom uniformly from the
// OccLocation list of atoms (one Lemire word, OccLocation.hh:257-261) and a
// direction (random_int(3): one word, its top two bits), is rejected when the
// target holds an atom, and otherwise is accepted without a draw when
// dE = J (a - b) <= 0 (a, b: atoms next to the origin other than at the
// target, next to the target other than at the origin), else when one more
// canonical word falls below exp(-beta dE) (a per-class word threshold from
// the host).  The accepted attempt is the selected event, with time increment
// fl(k / R) for the k attempts it took; the loop's time is the running FP64
// sum of the increments in event order.
//
// Device OccLocation with atom tracking (OccLocation.cc:39-118, 216-330): one
// 16-bit list array holds the vacancy list from the front and the atom list
// from the back (slot j at n - 1 - j); each site has its position in that
// array and the id of its atom (0xFFFF for none); each atom its displacement
// since the start and its jump count.  An event's two transforms (origin
// 1 -> 0, then target 0 -> 1) and the atom's move are six LDS writes.
//
// Block walk: lane t evaluates stream position P+t as if an attempt started
// there, from the block-start state.  Rejected attempts change nothing but the
// clock, and an accepted hop changes only what a later attempt reads at its
// own slots and sites, so the block commits every attempt up to the first one
// an earlier accepted hop changes and starts again after them.
#include "internal.hh"
#include "walk.hh"

namespace mcg {

namespace {

constexpr uint16_t NO_ATOM = 0xFFFFu;

__device__ __forceinline__ uint32_t kmc_nbr(uint32_t l, int d, int32_t rows, int32_t cols,
                                            uint64_t magic) {
  const uint32_t j = (uint32_t)(((uint64_t)l * magic) >> 40);
  const uint32_t i = l - j * (uint32_t)rows;
  switch (d) {
    case 0: return ((i + 1 == (uint32_t)rows) ? 0u : i + 1) + (uint32_t)rows * j;
    case 1: return ((i == 0) ? (uint32_t)rows - 1 : i - 1) + (uint32_t)rows * j;
    case 2: return i + (uint32_t)rows * ((j + 1 == (uint32_t)cols) ? 0u : j + 1);
    default: return i + (uint32_t)rows * ((j == 0) ? (uint32_t)cols - 1 : j - 1);
  }
}

// One attempt evaluated from a state: the atom's list slot and position, the
// origin and target, whether the target is free, the class a - b + 3, the
// atom and its displacement.
struct Attempt {
  int32_t pa;        // list position of the atom's slot
  uint32_t from, to;
  int dir;
  bool free;         // the target is a vacancy
  int cls;           // a - b + 3
  uint32_t atom;
  int32_t dx, dy;
};

__device__ __forceinline__ void kmc_eval(Attempt &at, uint32_t slot, int dir, int32_t n,
                                         const uint16_t *list, const uint16_t *site_atom,
                                         const int32_t *adx, const int32_t *ady, int32_t rows,
                                         int32_t cols, uint64_t magic) {
  at.pa = n - 1 - (int32_t)slot;
  at.from = list[at.pa];
  at.dir = dir;
  at.to = kmc_nbr(at.from, dir, rows, cols, magic);
  at.atom = site_atom[at.from];
  at.free = site_atom[at.to] == NO_ATOM;
  int a = 0, b = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t x = kmc_nbr(at.from, e, rows, cols, magic);
    const uint32_t y = kmc_nbr(at.to, e, rows, cols, magic);
    a += (x != at.to && site_atom[x] != NO_ATOM) ? 1 : 0;
    b += (y != at.from && site_atom[y] != NO_ATOM) ? 1 : 0;
  }
  at.cls = a - b + 3;
  at.dx = adx[at.atom];
  at.dy = ady[at.atom];
}

// OccLocation::apply for the two transforms (origin 1 -> 0, target 0 -> 1)
// and the atom's move, by one lane.  The atom list's last entry moves into the
// origin's slot, the origin takes the vacancy list's new last place, the
// vacancy list's last entry (now the origin) moves into the target's slot, and
// the target takes the atom list's new last place: the same position pba.
__device__ __forceinline__ void kmc_apply(const Attempt &at, int32_t pba, uint16_t *list,
                                          uint16_t *pos, uint16_t *site_atom, int32_t *adx,
                                          int32_t *ady, int32_t *ajump) {
  const int32_t pt = pos[at.to];
  const uint16_t back = list[pba];
  list[at.pa] = back;
  pos[back] = (uint16_t)at.pa;
  list[pt] = (uint16_t)at.from;
  pos[at.from] = (uint16_t)pt;
  list[pba] = (uint16_t)at.to;
  pos[at.to] = (uint16_t)pba;
  site_atom[at.from] = NO_ATOM;
  site_atom[at.to] = (uint16_t)at.atom;
  adx[at.atom] = at.dx + (at.dir == 0 ? 1 : at.dir == 1 ? -1 : 0);
  ady[at.atom] = at.dy + (at.dir == 2 ? 1 : at.dir == 3 ? -1 : 0);
  ajump[at.atom] += 1;
}

}  // namespace

__global__ __launch_bounds__(128) void kmc_kernel(const KmcParams *__restrict__ Pp,
                                                 ExactState *__restrict__ Sp,
                                                 int8_t *__restrict__ lat_pool,
                                                 int32_t *__restrict__ list_pool,
                                                 int32_t *__restrict__ atom_pool,
                                                 KmcHop *__restrict__ hop_pool,
                                                 int32_t *__restrict__ hop_count_pool,
                                                 mcg_sample *__restrict__ series,
                                                 int64_t series_cap,
                                                 const uint8_t *__restrict__ active,
                                                 int64_t n_passes, int64_t n_steps,
                                                 int64_t sample_period, double stop_time,
                                                 int select_at_end) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  const int lane = threadIdx.x & 63;
  const bool producer = threadIdx.x >= 64;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint64_t raw[3][MT_N];
  __shared__ uint64_t ring[3][MT_N];
  __shared__ int done;
  __shared__ uint64_t thr_s[KMC_NCLASS];

  const KmcParams *P = Pp + c;
  ExactState *St = Sp + c;
  const int32_t rows = P->rows, cols = P->cols, n = P->n, n_atoms = P->n_atoms;
  const int W = P->window;
  const uint64_t magic = P->div_magic;
  const double R = P->attempt_rate, period = P->time_period;
  int8_t *g_lat = lat_pool + P->lat_off;
  int32_t *g_list = list_pool + P->lat_off;
  int32_t *g_atoms = atom_pool + 4 * P->lat_off;  // per atom: site, dx, dy, jumps
  const int a_al = (n + 7) & ~7;
  uint16_t *list = reinterpret_cast<uint16_t *>(smem);
  uint16_t *pos = list + a_al;
  uint16_t *site_atom = pos + a_al;
  int32_t *adx = reinterpret_cast<int32_t *>(site_atom + a_al);
  int32_t *ady = adx + a_al;
  int32_t *ajump = ady + a_al;

  if (!producer) {
    for (int i = lane; i < MT_N; i += 64) {
      raw[0][i] = St->mt[i];
      ring[0][i] = mt_temper(raw[0][i]);
    }
    if (lane < KMC_NCLASS) thr_s[lane] = P->thr[lane];
    for (int i = lane; i < n; i += 64) {
      const uint16_t l = (uint16_t)g_list[i];
      list[i] = l;
      pos[l] = (uint16_t)i;
      site_atom[i] = NO_ATOM;
    }
    for (int i = lane; i < n_atoms; i += 64) {
      adx[i] = g_atoms[4 * i + 1];
      ady[i] = g_atoms[4 * i + 2];
      ajump[i] = g_atoms[4 * i + 3];
    }
    if (lane == 0) done = 0;
  }
  __syncthreads();
  if (!producer)
    for (int i = lane; i < n_atoms; i += 64) site_atom[g_atoms[4 * i]] = (uint16_t)i;
  __syncthreads();
  if (producer) {  // wave 1 twists the engine (walk.hh); wave 0 walks
    mt_producer(raw, ring, &done, lane);
    return;
  }

  const int64_t pos0 = St->pos;
  int64_t ps = pos0;
  MtRing3 eng;
  int64_t n_pass = St->n_pass, n_acc = St->n_accept, n_rej = St->n_reject;
  int64_t S = St->S, n_samples = St->n_samples;
  int64_t sum_sq = St->kmc_sum_sq, sum_dx = St->kmc_sum_dx, sum_dy = St->kmc_sum_dy;
  int64_t k_run = St->kmc_k;  // attempts since the last event
  double time = St->time;
  int64_t next_k = St->next_k;
  int64_t step = St->step;
  int64_t frozen = 0;
  const uint32_t draw_mask = P->draw_mask, full_mask = P->full_mask;
  const int64_t spin_limit = P->spin_limit;
  const int32_t pba = n - n_atoms;  // the atom list's last position (fixed: hops keep the count)
  const int64_t target = n_passes * (int64_t)n + n_steps;
  int64_t done_ev = 0;
  int64_t pending_site = St->pending_site, pending_dir = St->kmc_pdir, pending_k = St->kmc_pk;
  double pending_time = St->pending_dt;

  auto ensure = [&](int64_t upto) { eng.ensure(upto); };
  const uint64_t *ringf = &ring[0][0];
  auto word = [&](int64_t q) -> uint64_t { return ringf[eng.index(q)]; };
  auto lemire_scalar = [&](uint64_t m) -> uint32_t {
    ensure(ps + 1);
    uint64_t x = word(ps);
    ++ps;
    uint64_t lo = x * m, hi = __umul64hi(x, m);
    if (lo < m) {
      const uint64_t th = (0ULL - m) % m;
      while (lo < th) {
        ensure(ps + 1);
        x = word(ps);
        ++ps;
        lo = x * m;
        hi = __umul64hi(x, m);
      }
    }
    return (uint32_t)hi;
  };
  auto record = [&](int64_t pass_field, double t) {
    if (lane == 0 && n_samples < series_cap) {
      mcg_sample rec = {};
      rec.pass = pass_field;
      rec.S1 = S;
      rec.M = n_atoms;
      rec.S2 = sum_sq;
      rec.S3 = sum_dx;
      rec.T3 = sum_dy;
      rec.time = t;
      series[(int64_t)c * series_cap + n_samples] = rec;
    }
    ++n_samples;
  };
  // the executed event's bookkeeping (all lanes, uniform): time-grid samples
  // it reaches (before it applies), its time, bond and displacement sums
  auto account = [&](int64_t k_ev, int cls, int dir, int32_t dx, int32_t dy) {
    const double event_time = __dadd_rn(time, __ddiv_rn((double)k_ev, R));
    if (period > 0.0)
      while (event_time >= __dmul_rn((double)next_k, period)) {
        record(n_pass * (int64_t)n + step, __dmul_rn((double)next_k, period));
        ++next_k;
      }
    time = event_time;
    S += 3 - cls;  // b - a
    const int ddx = dir == 0 ? 1 : dir == 1 ? -1 : 0, ddy = dir == 2 ? 1 : dir == 3 ? -1 : 0;
    sum_dx += ddx;
    sum_dy += ddy;
    sum_sq += 2 * (int64_t)dx * ddx + 2 * (int64_t)dy * ddy + 1;
  };

  // hop correlations: after an atom's hop (by the lane that applied it, time
  // already the event's, step and pass still the event's own), its
  // displacement when its jump count is a multiple of hop_every
  const int32_t hop_every = P->hop_every, hop_max = P->hop_max;
  auto hop_record = [&](int32_t atom) {
    if (hop_every <= 0 || ajump[atom] % hop_every != 0) return;
    int32_t *cnt = hop_count_pool + P->lat_off;
    const int32_t k = cnt[atom];
    if (k >= hop_max) return;
    KmcHop r;
    r.dx = adx[atom];
    r.dy = ady[atom];
    r.step = step;
    r.pass = n_pass;
    r.time = time;
    hop_pool[P->hop_off + (int64_t)k * n + atom] = r;
    cnt[atom] = k + 1;
  };

  // a selected event: left pending (its site, direction, attempts and time
  // kept for the next launch) or applied
  bool stopped = false;
  auto pend = [&](uint32_t from, int dir, int64_t k, double event_time) {
    pending_site = from;
    pending_dir = dir;
    pending_k = k;
    pending_time = event_time;
    stopped = true;
  };
  auto bump = [&]() {
    ++n_acc;
    ++step;
    ++done_ev;
    if (step == n) {
      step = 0;
      ++n_pass;
      if (period <= 0.0 && n_pass % sample_period == 0) record(n_pass, time);
    }
  };

  // the event selected at the end of the previous launch comes first
  if (pending_site >= 0 && n_atoms > 0) {
    if (pending_time >= stop_time) {
      stopped = true;
    } else {
      Attempt a0;
      const uint32_t from = (uint32_t)pending_site;
      kmc_eval(a0, (uint32_t)(n - 1 - (int32_t)pos[from]), pending_dir, n, list, site_atom, adx,
               ady, rows, cols, magic);
      account(pending_k, a0.cls, a0.dir, a0.dx, a0.dy);
      if (lane == 0) {
        kmc_apply(a0, pba, list, pos, site_atom, adx, ady, ajump);
        hop_record(a0.atom);
      }
      asm volatile("" ::: "memory");
      pending_site = -1;
      bump();
    }
  }

  while (!stopped && n_atoms > 0) {
    // past the target: select one more event and leave it pending, if asked
    const bool select_only = done_ev >= target;
    if (select_only && !select_at_end) break;
    ensure(ps + W + 2);
    const bool inwin = lane < W;
    const int64_t q = inwin ? ps + lane : ps;
    const uint64_t x1 = word(q), x2 = word(q + 1), x3 = word(q + 2);
    const uint64_t na = (uint64_t)n_atoms;
    const uint32_t slot = (uint32_t)__umul64hi(x1, na);
    const bool special = inwin && (x1 * na) < na;
    Attempt at;
    kmc_eval(at, slot, (int)(x2 >> 62), n, list, site_atom, adx, ady, rows, cols, magic);
    // dE > 0 draws a word (draw_mask); all words may accept (full_mask)
    const bool nd = at.free && ((draw_mask >> at.cls) & 1u);
    const bool acc =
        at.free && (!((draw_mask >> at.cls) & 1u) || ((full_mask >> at.cls) & 1u) || x3 < thr_s[at.cls]);
    const bool valid = inwin && !special;
    const uint64_t ND = __ballot(valid && nd);
    const uint64_t AC = __ballot(valid && acc);
    const uint64_t SP = __ballot(special);

    // attempt starts over the window: 0, then +2 or +3 (a uniform drawn), up
    // to the first special one
    const uint64_t A = window_starts2(ND, SP, 0ull, W);
    // An attempt evaluated from the block-start state stays exact until an
    // earlier accepted hop changes what it reads: the hop rewrites its own
    // list slot and the atom list's last slot, and changes its origin and
    // target sites.  So an attempt is hit when its slot is one of those two,
    // or its origin, target or their neighbours include the hop's origin or
    // target (its class, its free target and its atom are read there).  The
    // block runs up to the first hit attempt; hits found from accepted lanes
    // past it change nothing.  A window without an accepted attempt skips all
    // of this.
    const uint64_t AA = A & AC;
    uint64_t exec = A;
    if (AA & (AA - 1ull) ? true : (AA && (A >> __builtin_ctzll(AA)) > 1ull)) {
      uint32_t nf[4], ntg[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        nf[e] = kmc_nbr(at.from, e, rows, cols, magic);
        ntg[e] = kmc_nbr(at.to, e, rows, cols, magic);
      }
      const bool is_att = (A >> lane) & 1ull;
      bool hit = false;
      uint64_t HIT = 0;
      for (uint64_t rem = AA; rem; rem &= rem - 1) {
        const int e = __builtin_ctzll(rem);
        if (HIT & ((1ull << e) - 1)) break;  // past the first hit attempt
        const int32_t pa_e = __builtin_amdgcn_readlane(at.pa, e);
        const uint32_t fe = (uint32_t)__builtin_amdgcn_readlane((int)at.from, e);
        const uint32_t te = (uint32_t)__builtin_amdgcn_readlane((int)at.to, e);
        bool h = at.pa == pa_e || at.pa == pba || at.from == fe || at.from == te ||
                 at.to == fe || at.to == te;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          h = h || nf[k] == fe || nf[k] == te || ntg[k] == fe || ntg[k] == te;
        hit = hit || (lane > e && h);
        HIT = __ballot(is_att && hit);
      }
      if (HIT) exec = A & ((1ull << __builtin_ctzll(HIT)) - 1);
    }

    // the attempts in order: rejected ones (counted in bulk between events)
    // add to the next event's increment; accepted ones are the events,
    // applied in order by their lanes (their slots and sites are disjoint from
    // every later attempt run here)
    int64_t kk = k_run;
    int last = -1;
    int attempts = 0;
    uint64_t consumed = 0;
    bool cut = false;
    for (uint64_t ev = exec & AC; ev; ev &= ev - 1) {
      const int t = __builtin_ctzll(ev);
      const uint64_t upto = t == 63 ? ~0ull : ((2ull << t) - 1ull);
      const uint64_t span = exec & upto & ~consumed;  // attempts since the last event, with t
      const int cnt = __popcll(span);
      consumed |= span;
      kk += cnt;
      attempts += cnt;
      n_rej += cnt - 1;
      last = t;
      const int cls = __builtin_amdgcn_readlane(at.cls, t);
      const int dir = __builtin_amdgcn_readlane(at.dir, t);
      const int32_t dx = __builtin_amdgcn_readlane(at.dx, t), dy = __builtin_amdgcn_readlane(at.dy, t);
      const uint32_t from = __builtin_amdgcn_readlane((int)at.from, t);
      const double event_time = __dadd_rn(time, __ddiv_rn((double)kk, R));
      if (done_ev >= target || event_time >= stop_time) {
        pend(from, dir, kk, event_time);
        kk = 0;
        cut = true;
        break;
      }
      account(kk, cls, dir, dx, dy);
      if (lane == t) {
        kmc_apply(at, pba, list, pos, site_atom, adx, ady, ajump);
        hop_record(at.atom);
      }
      asm volatile("" ::: "memory");
      bump();
      kk = 0;
      if (done_ev >= target && !select_at_end) {  // the launch's events are done
        cut = true;
        break;
      }
    }
    if (!cut) {  // rejected attempts after the last event
      const uint64_t rest = exec & ~consumed;
      const int cnt = __popcll(rest);
      kk += cnt;
      attempts += cnt;
      n_rej += cnt;
      if (rest) last = 63 - __builtin_clzll(rest);
    }
    k_run = kk;
    if (last >= 0) ps += last + 2 + (int)((ND >> last) & 1ull);

    if (!stopped && (SP & 1ull) && attempts == 0) {
      // an attempt whose Lemire draw may reject, with the full rejection loop;
      // every lane computes it, lane 0 writes
      const uint32_t s1 = lemire_scalar(na);
      ensure(ps + 1);
      const int d1 = (int)(word(ps) >> 62);
      ++ps;
      Attempt a1;
      kmc_eval(a1, s1, d1, n, list, site_atom, adx, ady, rows, cols, magic);
      bool ok = a1.free;
      if (a1.free && ((draw_mask >> a1.cls) & 1u)) {
        ensure(ps + 1);
        const uint64_t xr = word(ps);
        ++ps;
        ok = ((full_mask >> a1.cls) & 1u) || xr < thr_s[a1.cls];
      }
      ++k_run;
      // one wavefront walks: its LDS operations execute in issue order
      asm volatile("" ::: "memory");
      if (ok) {
        const double event_time = __dadd_rn(time, __ddiv_rn((double)k_run, R));
        if (select_only || event_time >= stop_time) {
          pend(a1.from, a1.dir, k_run, event_time);
        } else {
          account(k_run, a1.cls, a1.dir, a1.dx, a1.dy);
          if (lane == 0) {
            kmc_apply(a1, pba, list, pos, site_atom, adx, ady, ajump);
            hop_record(a1.atom);
          }
          asm volatile("" ::: "memory");
          bump();
        }
        k_run = 0;
      } else {
        ++n_rej;
      }
    }
    // no event in spin_limit attempts (every move of the state has a
    // vanishing probability): stop instead of spinning
    if (k_run > spin_limit) {
      frozen = 1;
      break;
    }
  }

  // release the producer (it waits at the barrier this one matches)
  eng.finish(&done, lane);
  int64_t g_fin = 0, p_fin = pos0;
  if (ps != pos0) {
    g_fin = (ps - 1) / MT_N;
    p_fin = ps - (int64_t)MT_N * g_fin;
  }
  for (int i = lane; i < MT_N; i += 64) St->mt[i] = raw[g_fin % 3][i];
  for (int i = lane; i < n; i += 64) {
    g_list[i] = (int32_t)list[i];
    g_lat[i] = site_atom[i] == NO_ATOM ? (int8_t)0 : (int8_t)1;
    const uint16_t a = site_atom[i];
    if (a != NO_ATOM) {
      g_atoms[4 * a + 0] = (int32_t)i;
      g_atoms[4 * a + 1] = adx[a];
      g_atoms[4 * a + 2] = ady[a];
      g_atoms[4 * a + 3] = ajump[a];
    }
  }
  if (lane == 0) {
    St->pos = p_fin;
    St->n_pass = n_pass;
    St->n_accept = n_acc;
    St->n_reject = n_rej;
    St->S = S;
    St->M = n_atoms;
    St->n_samples = n_samples;
    St->step = step;
    St->time = time;
    St->next_k = next_k;
    St->kmc_k = k_run;
    St->kmc_sum_sq = sum_sq;
    St->kmc_sum_dx = sum_dx;
    St->kmc_sum_dy = sum_dy;
    St->frozen = frozen;
    St->pending_site = pending_site;
    St->pending_dt = pending_time;
    St->kmc_pdir = pending_dir;
    St->kmc_pk = pending_k;
  }
}

void launch_kmc(const KmcParams *P, ExactState *S, int8_t *lat_pool, int32_t *list_pool,
                int32_t *atom_pool, KmcHop *hop_pool, int32_t *hop_count_pool,
                mcg_sample *series, int64_t series_cap, const uint8_t *active, int n_chains,
                int64_t n_passes, int64_t n_steps, int64_t sample_period, double stop_time,
                int select_at_end, int lds_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(kmc_kernel, dim3(n_chains), dim3(128), lds_bytes, stream, P, S, lat_pool,
                     list_pool, atom_pool, hop_pool, hop_count_pool, series, series_cap, active,
                     n_passes, n_steps, sample_period, stop_time, select_at_end);
}

hipError_t kmc_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kmc_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
