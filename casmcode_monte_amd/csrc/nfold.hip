// nfold.hip -- NFOLD mode: the N-fold way (rejection-free kinetic Monte Carlo)
// for the SGC square-lattice Ising model, one sequential chain per wavefront.
//
// Reference: the main loop of nfold.hh:78-149 (select an event and its time
// increment, sample by count with the increment as weight or on the time grid,
// apply, advance time, count a step).  The event selector is defined here:
// a site's rate is its Metropolis probability for the SGC potential delta of
// its class k = (s > 0) * 5 + (nbr_sum + 4) / 2 (the EXACT mode's classes,
// basic_semigrand_canonical.hh:178-192, model.hh:305-345); sites are kept in
// one array partitioned by class, class k at [start[k], start[k+1]).  Per
// event: R = sum_k n_k r_k in class order; the class is the first whose
// running sum exceeds u1 * R; the site is arr[start[k] + p] with p a Lemire
// draw over n_k; dt = -log1p(-u3) / R.  After the flip the site and its four
// neighbours (+i, -i, +j, -j, in that order) move to their new classes: down
// a class by swapping with the first entry of the segment left, up by
// swapping with the last.
//
// Layout: workgroup = one 64-lane wavefront = one chain.  LDS holds the class
// array and each site's position in it (16-bit each; bit 15 of a position is
// the site's spin, which its class also fixes: 4 bytes a site, so two 128x128
// chains share a CU) and the MT19937-64 ring.  The segment starts live in
// registers (lane k holds start[k]); the rates likewise (lane k, k < 10).
// The walk is sequential; its parallel parts: the running rate sums and the
// class choice (ballot over lanes 0..9), the five sites' neighbourhoods (lanes
// 0..4, one LDS round trip), each class move as one rotation (one lane per
// segment boundary crossed), and the time increments of 64 events at once.
#include "internal.hh"
#include "walk.hh"

namespace mcg {

namespace {

// pos[l] holds l's position in the class array (bits 0..14) and its spin
// (bit 15: 1 = +1)
constexpr uint16_t SPIN_BIT = 0x8000u;
__device__ __forceinline__ int spin_at(const uint16_t *pos, uint32_t l) {
  return (pos[l] & SPIN_BIT) ? 1 : -1;
}

__device__ __forceinline__ void nbrs(uint32_t l, int32_t rows, int32_t cols, uint64_t magic,
                                     uint32_t out[4]) {
  const uint32_t j = (uint32_t)(((uint64_t)l * magic) >> 40);
  const uint32_t i = l - j * (uint32_t)rows;
  const uint32_t iu = (i + 1 == (uint32_t)rows) ? 0u : i + 1;
  const uint32_t id = (i == 0) ? (uint32_t)rows - 1 : i - 1;
  const uint32_t ju = (j + 1 == (uint32_t)cols) ? 0u : j + 1;
  const uint32_t jd = (j == 0) ? (uint32_t)cols - 1 : j - 1;
  out[0] = iu + (uint32_t)rows * j;
  out[1] = id + (uint32_t)rows * j;
  out[2] = i + (uint32_t)rows * ju;
  out[3] = i + (uint32_t)rows * jd;
}

__device__ __forceinline__ int class_of(const uint16_t *pos, uint32_t l, int32_t rows,
                                        int32_t cols, uint64_t magic) {
  uint32_t nb[4];
  nbrs(l, rows, cols, magic, nb);
  const int nsum = spin_at(pos, nb[0]) + spin_at(pos, nb[1]) + spin_at(pos, nb[2]) +
                   spin_at(pos, nb[3]);
  return (spin_at(pos, l) > 0 ? 5 : 0) + ((nsum + 4) >> 1);
}

// A cross-lane read only sees lanes that are active when it executes.  The
// empty asm consumes the result where the call stands (all 64 lanes active),
// so the compiler cannot sink the read into a branch that masks off its
// source lanes.
template <typename T>
__device__ __forceinline__ T pinned(T v) {
  asm volatile("" : "+v"(v));
  return v;
}
// lane k receives lane k+1's value (row_shl:1) / lane k-1's (row_shr:1)
// within its row of 16 lanes; the walk uses lanes 0..10
__device__ __forceinline__ int row_shl1(int v) {
  return pinned(__builtin_amdgcn_update_dpp(0, v, 0x101, 0xf, 0xf, false));
}
__device__ __forceinline__ int row_shr1(int v) {
  return pinned(__builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));
}

}  // namespace

__global__ __launch_bounds__(64) void nfold_kernel(const NfoldParams *__restrict__ Pp,
                                                   ExactState *__restrict__ Sp,
                                                   int8_t *__restrict__ lat_pool,
                                                   int32_t *__restrict__ list_pool,
                                                   mcg_sample *__restrict__ series,
                                                   int64_t series_cap,
                                                   const uint8_t *__restrict__ active,
                                                   int64_t n_passes, int64_t sample_period,
                                                   int sample_now, int64_t max_events,
                                                   double stop_time, int select_at_end) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  const int lane = threadIdx.x;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint64_t raw[2][MT_N];
  __shared__ uint64_t ring[RING];
  __shared__ int hist[16];

  const NfoldParams *P = Pp + c;
  ExactState *St = Sp + c;
  const int32_t rows = P->rows, cols = P->cols, n = P->n;
  const uint64_t magic = P->div_magic;
  const double period = P->time_period;
  const bool by_time = period > 0.0;
  int8_t *g_lat = lat_pool + P->lat_off;
  int32_t *g_arr = list_pool + P->lat_off;
  uint16_t *arr = reinterpret_cast<uint16_t *>(smem);
  uint16_t *pos = arr + n;

  for (int i = lane; i < MT_N; i += 64) {
    raw[0][i] = St->mt[i];
    ring[i] = mt_temper(raw[0][i]);
  }
  if (lane < 16) hist[lane] = 0;
  for (int i = lane; i < n; i += 64) {
    const uint32_t l = (uint32_t)g_arr[i];
    arr[i] = (uint16_t)l;
    pos[l] = (uint16_t)(i | (g_lat[l] > 0 ? SPIN_BIT : 0u));
  }
  __syncthreads();
  for (int l = lane; l < n; l += 64) atomicAdd(&hist[class_of(pos, l, rows, cols, magic)], 1);
  __syncthreads();
  // lane k holds start[k] (k <= 10) and rate[k] (k < 10)
  int startv = 0;
  for (int k = 0; k < 10; ++k)
    if (k < lane) startv += hist[k];
  const double ratev = lane < 10 ? P->rate[lane] : 0.0;

  const int64_t pos0 = St->pos;
  int64_t pos_mt = pos0;
  int64_t G = 0;
  int64_t n_pass = St->n_pass, n_ev = St->n_accept, n_samples = St->n_samples;
  int64_t S = St->S, M = St->M, next_k = St->next_k;
  double time = St->time;
  int64_t pending = St->pending_site;
  double pending_dt = St->pending_dt;
  double pending_R = St->e_run;  // the total rate at the pending event's selection
  int64_t step = St->step;       // events done in the current pass
  const int self_nb = 2 * (rows == 1) + 2 * (cols == 1);
  const bool narrow = rows <= 2 || cols <= 2;  // a neighbour can be listed twice
  bool frozen = false;  // total rate 0: the next event never comes

  auto ensure = [&](int64_t upto) {
    while (upto > (int64_t)MT_N * (G + 1)) {
      twist_gen(raw[G & 1], raw[(G + 1) & 1], ring + ((G + 1) & 1) * MT_N, lane);
      ++G;
    }
  };
  // ring slot of stream position q, which ensure() has made available
  auto ring_slot = [&](int64_t q) -> int {
    const int64_t base = (int64_t)MT_N * (G - 1);
    int r = (int)(q - base) + (int)(((G - 1) & 1) * MT_N);
    return r >= RING ? r - RING : r;
  };
  auto ring_at = [&](int64_t q) -> uint64_t { return rfl64(ring[ring_slot(q)]); };
  auto word = [&]() -> uint64_t {
    ensure(pos_mt + 1);
    const uint64_t w = ring_at(pos_mt);
    ++pos_mt;
    return w;
  };
  // select an event: the site l, its position, the time-increment word u3 and
  // the total rate R.  Lane k (k < 10) holds n_k and its running rate sum; the
  // sum is formed in class order, so running sums equal the sequential ones
  // (an empty class adds +0.0).
  auto select = [&](uint32_t &site, int &site_pos, double &u3, double &R) {
    // the usual three words go out first and arrive while the rates are summed
    ensure(pos_mt + 3);
    const uint64_t v0 = ring[ring_slot(pos_mt)], v1 = ring[ring_slot(pos_mt + 1)],
                   v2 = ring[ring_slot(pos_mt + 2)];
    const int nxt = row_shl1(startv);
    const int nv = lane < 10 ? nxt - startv : 0;
    const double tv = __dmul_rn((double)nv, ratev);
    double run = 0.0, pre = 0.0;
#pragma unroll
    for (int k = 0; k < 10; ++k) {
      run = __dadd_rn(run, readlane_f64(tv, k));
      pre = lane == k ? run : pre;
    }
    R = run;
    const uint64_t w0 = rfl64(v0), w1 = rfl64(v1), w2 = rfl64(v2);
    pos_mt += 2;
    const double target = __dmul_rn(canonical_from_word(w0), R);
    const uint64_t hit = __ballot(lane < 10 && nv > 0 && target < pre);
    int kk;
    if (hit) {
      kk = __ffsll((unsigned long long)hit) - 1;
    } else {  // rounding left none: the last non-empty class
      kk = 63 - __clzll((unsigned long long)__ballot(lane < 10 && nv > 0));
    }
    const uint64_t nk = (uint64_t)__builtin_amdgcn_readlane(nv, kk);
    uint64_t x = w1, lo = x * nk, hi = __umul64hi(x, nk);
    bool redrawn = false;
    if (lo < nk) {
      const uint64_t th = (0ULL - nk) % nk;
      while (lo < th) {
        x = word();
        lo = x * nk;
        hi = __umul64hi(x, nk);
        redrawn = true;
      }
    }
    uint64_t w3 = w2;
    if (redrawn)
      w3 = word();
    else
      ++pos_mt;
    u3 = canonical_from_word(w3);
    site_pos = __builtin_amdgcn_readlane(startv, kk) + (int)hi;
    site = rfl32((int32_t)arr[site_pos]);
  };
  // Move site x from class a to class b at position p.  Going up, x swaps in
  // turn with the last entry of segments a, a+1, .., b-1 (each of which then
  // starts one lower); going down, with the first entry of segments a, a-1,
  // .., b+1 (each then starts one higher).  The chain of swaps is a rotation:
  // lane k holds the boundary between segments k-1 and k, reads the entry
  // there and writes it into the previous hole, all lanes at once; a boundary
  // that coincides with the previous one (empty segment) is a swap of x with
  // itself and writes nothing.
  auto move = [&](uint32_t x, int a, int b, int p_known) {
    const bool up = a < b;
    // lane k: the last entry of segment k-1 (up) or the first of segment k (down)
    const int q = up ? startv - 1 : startv;
    const int from_below = row_shr1(q), from_above = row_shl1(q);
    const bool act = up ? (lane > a && lane <= b) : (lane > b && lane <= a);
    uint32_t y = arr[act ? q : 0], px = pos[x];
    asm volatile("" : "+v"(y), "+v"(px));  // both reads issued before either is used
    const int p = p_known >= 0 ? p_known : rfl32((int32_t)(px & 0x7fffu));
    int qp = lane == (up ? a + 1 : a) ? p : up ? from_below : from_above;
    const int qx = __builtin_amdgcn_readlane(q, up ? b : b + 1);
    // an entry keeps its class (k - 1 going up, k going down), which fixes its
    // spin; lane 11 (outside the segment lanes) carries x to its final slot,
    // in class b, in the same writes
    int cls = up ? lane - 1 : lane;
    bool wr = act && q != qp;
    if (lane == 11) {
      y = x;
      qp = qx;
      cls = b;
      wr = true;
    }
    if (wr) {
      arr[qp] = (uint16_t)y;
      pos[y] = (uint16_t)(qp | (cls >= 5 ? SPIN_BIT : 0u));
    }
    startv += act ? (up ? -1 : 1) : 0;
    // the wave's LDS operations execute in issue order; keep the compiler from
    // moving the next move's reads above these writes
    asm volatile("" ::: "memory");
  };
  // a by-pass record also keeps the total rate R in e_run
  auto record = [&](double t, double w, double R) {
    if (lane == 0 && n_samples < series_cap) {
      mcg_sample rec = {};
      rec.pass = n_pass;
      rec.S1 = S;
      rec.M = M;
      rec.e_run = R;
      rec.time = t;
      rec.weight = w;
      series[(int64_t)c * series_cap + n_samples] = rec;
    }
    ++n_samples;
  };
  // by pass: select the pending event and sample the state with its increment
  auto sample_pass = [&]() {
    uint32_t l;
    int l_pos;
    double u3, R;
    select(l, l_pos, u3, R);
    pending_dt = -log1p(-u3) / R;
    pending = l;
    pending_R = R;
    record(time, pending_dt, R);
  };
  // the pass-0 sample, or a sample of the current pass asked for by the host,
  // unless this pass's first event is already selected
  if (max_events < 0 && !by_time && pending < 0 &&
      (sample_now || (P->sample_start && n_pass == 0 && n_samples == 0)))
    sample_pass();

  // Deferred time increments: lane j holds event j of the current batch (its
  // u3 and R, or its known increment with R = -1, and the state before it).
  // A flush turns them into increments on all lanes at once, adds them to the
  // time in event order and, on the time grid, writes the batch's records.
  double bu = 0.0, bR = 1.0;
  int64_t bS = 0, bM = 0;
  int nb_ev = 0;
  auto flush = [&]() {
    if (nb_ev == 0) return;
    const double dtv = bR < 0.0 ? bu : -log1p(-bu) / bR;
    double tend = 0.0;
#pragma unroll
    for (int j = 0; j < 64; ++j) {  // constant lane indices: no lane-select hazards
      if (j < nb_ev) {
        time = __dadd_rn(time, readlane_f64(dtv, j));
        tend = lane == j ? time : tend;
      }
    }
    if (by_time) {
      // first grid index k >= next_k with k * period >= t_end (the predicate
      // decides at the edge); t_end is non-decreasing, so each lane can
      // search from the batch's starting index
      int64_t k = next_k;
      const bool never = lane < nb_ev && !(tend <= 1.79769313486231570815e308);
      if (__ballot(never)) frozen = true;
      if (lane < nb_ev && !never) {
        k = (int64_t)ceil(tend / period);
        if (k < next_k) k = next_k;
        while (k > next_k && (double)(k - 1) * period >= tend) --k;
        while ((double)k * period < tend) ++k;
      }
      const int64_t ku = pinned(__shfl_up(k, 1));
      const int64_t kp = lane == 0 ? next_k : ku;
      const int64_t m = lane < nb_ev ? k - kp : 0;
      const uint64_t rec = __ballot(m > 0);
      if (m > 0) {
        const int64_t ns =
            n_samples + __popcll((unsigned long long)rec & ((1ULL << lane) - 1ULL));
        if (ns < series_cap) {
          mcg_sample r = {};
          r.pass = n_pass;
          r.S1 = bS;
          r.M = bM;
          r.time = (double)kp * period;
          r.weight = (double)m;
          series[(int64_t)c * series_cap + ns] = r;
        }
      }
      n_samples += __popcll((unsigned long long)rec);
      const int32_t klo = __builtin_amdgcn_readlane((int32_t)k, nb_ev - 1);
      const int32_t khi = __builtin_amdgcn_readlane((int32_t)(k >> 32), nb_ev - 1);
      next_k = ((int64_t)khi << 32) | (uint32_t)klo;
    }
    nb_ev = 0;
  };

  // flip site l (at position l_pos when known, else -1), update S and M, and
  // re-class it and its four neighbours
  auto apply = [&](uint32_t l, int l_pos) {
    // lanes 0..4 take l, +i, -i, +j, -j: their spin, neighbour sum, and how
    // many of their neighbour slots hold l
    const uint32_t lj = (uint32_t)(((uint64_t)l * magic) >> 40);
    const uint32_t li = l - lj * (uint32_t)rows;
    uint32_t xi = li, xj = lj;
    if (lane == 1) xi = li + 1 == (uint32_t)rows ? 0u : li + 1;
    if (lane == 2) xi = li == 0 ? (uint32_t)rows - 1 : li - 1;
    if (lane == 3) xj = lj + 1 == (uint32_t)cols ? 0u : lj + 1;
    if (lane == 4) xj = lj == 0 ? (uint32_t)cols - 1 : lj - 1;
    const uint32_t xs = xi + (uint32_t)rows * xj;
    const uint32_t xu = xi + 1 == (uint32_t)rows ? 0u : xi + 1;
    const uint32_t xd = xi == 0 ? (uint32_t)rows - 1 : xi - 1;
    const uint32_t yu = xj + 1 == (uint32_t)cols ? 0u : xj + 1;
    const uint32_t yd = xj == 0 ? (uint32_t)cols - 1 : xj - 1;
    const uint32_t xn[4] = {xu + (uint32_t)rows * xj, xd + (uint32_t)rows * xj,
                            xi + (uint32_t)rows * yu, xi + (uint32_t)rows * yd};
    const int sx = spin_at(pos, xs);
    const int nsx = spin_at(pos, xn[0]) + spin_at(pos, xn[1]) + spin_at(pos, xn[2]) +
                    spin_at(pos, xn[3]);
    const int mx = (xn[0] == l) + (xn[1] == l) + (xn[2] == l) + (xn[3] == l);
    const int s_old = __builtin_amdgcn_readlane(sx, 0);
    const int nsum = __builtin_amdgcn_readlane(nsx, 0);
    const int c_pre = (sx > 0 ? 5 : 0) + ((nsx + 4) >> 1);
    const int sp = xs == l ? -sx : sx;
    const int nsp = nsx - 2 * s_old * mx;
    const int c_post = (sp > 0 ? 5 : 0) + ((nsp + 4) >> 1);
    S += -2LL * s_old * (nsum - self_nb * s_old);
    M += -2LL * s_old;
    // the flip itself is written with l's move (its spin bit follows its class)
    for (int t = 0; t < 5; ++t) {
      const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)xs, t);
      const int a = __builtin_amdgcn_readlane(c_pre, t), b = __builtin_amdgcn_readlane(c_post, t);
      bool dup = false;  // listed twice (a lattice 1 or 2 sites wide): moved once
      if (narrow)
        for (int u = 0; u < t; ++u) dup |= (uint32_t)__builtin_amdgcn_readlane((int)xs, u) == x;
      if (dup || a == b) continue;
      move(x, a, b, t == 0 ? l_pos : -1);
    }
  };

  // Event by event (max_events >= 0; no records: the host samples), the
  // select-then-sample protocol of nfold.hh:121-141: the pending event is
  // applied first unless its time reaches stop_time; then events are selected
  // and, each time increment formed at once, applied until max_events are
  // done; an event whose time reaches stop_time, or (select_at_end) the one
  // after the last, is selected and left pending with its increment and the
  // total rate at its selection.  A total rate of 0 (no event ever comes)
  // stops the chain as frozen.
  if (max_events >= 0) {
    int64_t done = 0;
    bool stopped = false;
    auto bump = [&]() {
      ++n_ev;
      ++done;
      if (++step == n) {
        step = 0;
        ++n_pass;
      }
    };
    if (pending >= 0) {
      const double et = __dadd_rn(time, pending_dt);
      if (et >= stop_time) {
        stopped = true;
      } else {
        apply((uint32_t)pending, -1);
        time = et;
        pending = -1;
        bump();
      }
    }
    while (!stopped) {
      const bool select_only = done >= max_events;
      if (select_only && !select_at_end) break;
      uint32_t l;
      int l_pos;
      double u3, R;
      select(l, l_pos, u3, R);
      const double dt = -log1p(-u3) / R;
      const double et = __dadd_rn(time, dt);
      if (!(et <= 1.79769313486231570815e308)) {
        frozen = true;
        break;
      }
      if (select_only || et >= stop_time) {
        pending = l;
        pending_dt = dt;
        pending_R = R;
        stopped = true;
      } else {
        apply(l, l_pos);
        time = et;
        bump();
      }
    }
    n_passes = 0;
  }

  for (int64_t pp = 0; pp < n_passes; ++pp) {
    // a pass begun by event-by-event advances ends at its boundary
    const int32_t st0 = pp == 0 ? (int32_t)step : 0;
    step = 0;
    for (int32_t st = st0; st < n; ++st) {
      uint32_t l;
      int l_pos = -1;
      double u3, R;
      if (pending >= 0) {
        l = (uint32_t)pending;
        u3 = pending_dt;
        R = -1.0;
        pending = -1;
      } else {
        select(l, l_pos, u3, R);
      }
      if (lane == nb_ev) {
        bu = u3;
        bR = R;
        bS = S;
        bM = M;
      }
      ++nb_ev;
      apply(l, l_pos);
      ++n_ev;
      if (nb_ev == 64) flush();
      if (frozen) break;
    }
    flush();
    if (frozen) break;
    ++n_pass;
    if (!by_time && n_pass % sample_period == 0) sample_pass();
  }

  __syncthreads();
  int64_t g_fin = 0, p_fin = pos0;
  if (pos_mt != pos0) {
    g_fin = (pos_mt - 1) / MT_N;
    p_fin = pos_mt - (int64_t)MT_N * g_fin;
  }
  for (int i = lane; i < MT_N; i += 64) St->mt[i] = raw[g_fin & 1][i];
  for (int i = lane; i < n; i += 64) {
    g_lat[i] = (pos[i] & SPIN_BIT) ? (int8_t)1 : (int8_t)-1;
    g_arr[i] = (int32_t)arr[i];
  }
  if (lane == 0) {
    St->pos = p_fin;
    St->n_pass = n_pass;
    St->n_accept = n_ev;
    St->S = S;
    St->M = M;
    St->n_samples = n_samples;
    St->time = time;
    St->pending_site = pending;
    St->pending_dt = pending_dt;
    St->e_run = pending_R;
    St->step = step;
    St->next_k = next_k;
    St->frozen = frozen ? 1 : 0;
  }
}

void launch_nfold(const NfoldParams *P, ExactState *S, int8_t *lat_pool, int32_t *list_pool,
                  mcg_sample *series, int64_t series_cap, const uint8_t *active, int n_chains,
                  int64_t n_passes, int64_t sample_period, int sample_now, int64_t max_events,
                  double stop_time, int select_at_end, int lds_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(nfold_kernel, dim3(n_chains), dim3(64), lds_bytes, stream, P, S, lat_pool,
                     list_pool, series, series_cap, active, n_passes, sample_period, sample_now,
                     max_events, stop_time, select_at_end);
}

hipError_t nfold_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(nfold_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
