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
// Layout: workgroup = one 64-lane wavefront = one chain.  LDS holds the
// lattice (one byte per site, 1 = spin +1), the class array and each site's
// position in it (16-bit), and the MT19937-64 ring.  The segment starts live
// in registers (lane k holds start[k]); the rates likewise (lane k, k < 10).
// The walk is sequential: all lanes execute it with uniform operands; lanes
// 0..4 recompute the five classes after a flip in parallel.
#include "internal.hh"
#include "walk.hh"

namespace mcg {

namespace {

__device__ __forceinline__ int spin_at(const uint8_t *lat, uint32_t l) { return lat[l] ? 1 : -1; }

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

__device__ __forceinline__ int class_of(const uint8_t *lat, uint32_t l, int32_t rows,
                                        int32_t cols, uint64_t magic) {
  uint32_t nb[4];
  nbrs(l, rows, cols, magic, nb);
  const int nsum = spin_at(lat, nb[0]) + spin_at(lat, nb[1]) + spin_at(lat, nb[2]) +
                   spin_at(lat, nb[3]);
  return (lat[l] ? 5 : 0) + ((nsum + 4) >> 1);
}

}  // namespace

__global__ __launch_bounds__(64) void nfold_kernel(const NfoldParams *__restrict__ Pp,
                                                   ExactState *__restrict__ Sp,
                                                   int8_t *__restrict__ lat_pool,
                                                   int32_t *__restrict__ list_pool,
                                                   mcg_sample *__restrict__ series,
                                                   int64_t series_cap,
                                                   const uint8_t *__restrict__ active,
                                                   int64_t n_passes, int64_t sample_period) {
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
  uint8_t *lat = smem;
  uint16_t *arr = reinterpret_cast<uint16_t *>(smem + ((n + 15) & ~15));
  uint16_t *pos = arr + n;

  for (int i = lane; i < MT_N; i += 64) {
    raw[0][i] = St->mt[i];
    ring[i] = mt_temper(raw[0][i]);
  }
  if (lane < 16) hist[lane] = 0;
  for (int i = lane; i < n; i += 64) lat[i] = g_lat[i] > 0 ? 1u : 0u;
  __syncthreads();
  for (int i = lane; i < n; i += 64) {
    const uint32_t l = (uint32_t)g_arr[i];
    arr[i] = (uint16_t)l;
    pos[l] = (uint16_t)i;
    atomicAdd(&hist[class_of(lat, l, rows, cols, magic)], 1);
  }
  __syncthreads();
  // lane k holds start[k] (k <= 10) and rate[k] (k < 10)
  int startv = 0;
  for (int k = 0; k < 10; ++k)
    if (k < lane) startv += hist[k];
  const double ratev = lane < 10 ? P->rate[lane] : 0.0;
  auto start_of = [&](int k) -> int { return __builtin_amdgcn_readlane(startv, k); };

  const int64_t pos0 = St->pos;
  int64_t pos_mt = pos0;
  int64_t G = 0;
  int64_t n_pass = St->n_pass, n_ev = St->n_accept, n_samples = St->n_samples;
  int64_t S = St->S, M = St->M, next_k = St->next_k;
  double time = St->time;
  int64_t pending = St->pending_site;
  double pending_dt = St->pending_dt;

  auto ensure = [&](int64_t upto) {
    while (upto > (int64_t)MT_N * (G + 1)) {
      twist_gen(raw[G & 1], raw[(G + 1) & 1], ring + ((G + 1) & 1) * MT_N, lane);
      ++G;
    }
  };
  auto word = [&]() -> uint64_t {
    ensure(pos_mt + 1);
    const int64_t base = (int64_t)MT_N * (G - 1);
    int r = (int)(pos_mt - base) + (int)(((G - 1) & 1) * MT_N);
    if (r >= RING) r -= RING;
    ++pos_mt;
    return rfl64(ring[r]);
  };
  // select an event: site and time increment (three words and Lemire redraws)
  auto select = [&](uint32_t &site) -> double {
    double R = 0.0;
    for (int k = 0; k < 10; ++k)
      R = __dadd_rn(R, __dmul_rn((double)(start_of(k + 1) - start_of(k)), readlane_f64(ratev, k)));
    const double target = __dmul_rn(canonical_from_word(word()), R);
    double run = 0.0;
    int kk = -1, last = -1;
    for (int q = 0; q < 10; ++q) {
      const int nq = start_of(q + 1) - start_of(q);
      if (nq == 0) continue;
      last = q;
      run = __dadd_rn(run, __dmul_rn((double)nq, readlane_f64(ratev, q)));
      if (target < run) {
        kk = q;
        break;
      }
    }
    if (kk < 0) kk = last;
    const uint64_t nk = (uint64_t)(start_of(kk + 1) - start_of(kk));
    uint64_t x = word();
    uint64_t lo = x * nk, hi = __umul64hi(x, nk);
    if (lo < nk) {
      const uint64_t th = (0ULL - nk) % nk;
      while (lo < th) {
        x = word();
        lo = x * nk;
        hi = __umul64hi(x, nk);
      }
    }
    site = rfl32((int32_t)arr[start_of(kk) + (int)hi]);
    const double u3 = canonical_from_word(word());
    return -log1p(-u3) / R;
  };
  // move site x (now at position p) from class a to class b
  auto move = [&](uint32_t x, int b) {
    int p = rfl32((int32_t)pos[x]);
    int a = 0;
    for (int k = 1; k <= 10; ++k) a += start_of(k) <= p ? 1 : 0;
    while (a < b) {  // become the first entry of segment a+1
      const int q = start_of(a + 1) - 1;
      const uint32_t y = rfl32((int32_t)arr[q]);
      if (lane == 0) {
        arr[q] = (uint16_t)x;
        arr[p] = (uint16_t)y;
        pos[y] = (uint16_t)p;
      }
      p = q;
      if (lane == a + 1) --startv;
      ++a;
    }
    while (a > b) {  // become the last entry of segment a-1
      const int q = start_of(a);
      const uint32_t y = rfl32((int32_t)arr[q]);
      if (lane == 0) {
        arr[q] = (uint16_t)x;
        arr[p] = (uint16_t)y;
        pos[y] = (uint16_t)p;
      }
      p = q;
      if (lane == a) ++startv;
      --a;
    }
    if (lane == 0) pos[x] = (uint16_t)p;
    // the wave's LDS operations execute in issue order; keep the compiler from
    // moving the next move's reads above these writes
    asm volatile("" ::: "memory");
  };
  auto record = [&](double t, double w) {
    if (lane == 0 && n_samples < series_cap) {
      mcg_sample rec = {};
      rec.pass = n_pass;
      rec.S1 = S;
      rec.M = M;
      rec.time = t;
      rec.weight = w;
      series[(int64_t)c * series_cap + n_samples] = rec;
    }
    ++n_samples;
  };

  for (int64_t pp = 0; pp < n_passes; ++pp) {
    for (int32_t st = 0; st < n; ++st) {
      uint32_t l;
      double dt;
      if (pending >= 0) {
        l = (uint32_t)pending;
        dt = pending_dt;
        pending = -1;
      } else {
        dt = select(l);
      }
      if (by_time) {  // grid points inside [time, time + dt) see the current state
        // first grid index k with k * period >= t_end (the predicate decides at the edge)
        const double t_end = time + dt;
        int64_t k = (int64_t)ceil(t_end / period);
        if (k < next_k) k = next_k;
        while (k > next_k && (double)(k - 1) * period >= t_end) --k;
        while ((double)k * period < t_end) ++k;
        const int64_t m = k - next_k;
        if (m > 0) {
          record((double)next_k * period, (double)m);
          next_k += m;
        }
      }
      // flip, observables, then the five classes
      uint32_t nb[4];
      nbrs(l, rows, cols, magic, nb);
      const int s_old = spin_at(lat, l);
      const int nsum = spin_at(lat, nb[0]) + spin_at(lat, nb[1]) + spin_at(lat, nb[2]) +
                       spin_at(lat, nb[3]);
      const int self_nb = 2 * (rows == 1) + 2 * (cols == 1);
      S += -2LL * s_old * (nsum - self_nb * s_old);
      M += -2LL * s_old;
      __syncthreads();
      if (lane == 0) lat[l] ^= 1u;
      __syncthreads();
      const uint32_t xs = lane == 1 ? nb[0] : lane == 2 ? nb[1] : lane == 3 ? nb[2]
                          : lane == 4 ? nb[3] : l;
      const int xc = lane < 5 ? class_of(lat, xs, rows, cols, magic) : 0;
      for (int t = 0; t < 5; ++t)
        move((uint32_t)__builtin_amdgcn_readlane((int)xs, t), __builtin_amdgcn_readlane(xc, t));
      time += dt;
      ++n_ev;
    }
    ++n_pass;
    if (!by_time && n_pass % sample_period == 0) {
      uint32_t l;
      pending_dt = select(l);
      pending = l;
      record(time, pending_dt);
    }
  }

  __syncthreads();
  int64_t g_fin = 0, p_fin = pos0;
  if (pos_mt != pos0) {
    g_fin = (pos_mt - 1) / MT_N;
    p_fin = pos_mt - (int64_t)MT_N * g_fin;
  }
  for (int i = lane; i < MT_N; i += 64) St->mt[i] = raw[g_fin & 1][i];
  for (int i = lane; i < n; i += 64) {
    g_lat[i] = lat[i] ? (int8_t)1 : (int8_t)-1;
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
    St->next_k = next_k;
  }
}

void launch_nfold(const NfoldParams *P, ExactState *S, int8_t *lat_pool, int32_t *list_pool,
                  mcg_sample *series, int64_t series_cap, const uint8_t *active, int n_chains,
                  int64_t n_passes, int64_t sample_period, int lds_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(nfold_kernel, dim3(n_chains), dim3(64), lds_bytes, stream, P, S, lat_pool,
                     list_pool, series, series_cap, active, n_passes, sample_period);
}

hipError_t nfold_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(nfold_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
