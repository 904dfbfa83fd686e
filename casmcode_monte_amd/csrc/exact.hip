// exact.hip -- EXACT mode: the reference's sequential random-site Metropolis
// chain, reproduced bit for bit on one wavefront per chain.
//
// Reference: the main loop of basic_occupation_metropolis.hh:381-422 driving the
// SGC Ising calculator (basic_semigrand_canonical.hh:308-320 propose/apply,
// :178-192 potential delta, model.hh:305-345 formation-energy delta) with
// metropolis_acceptance (metropolis.hh:25-35) and RandomNumberGenerator over
// std::mt19937_64 (RandomNumberGenerator.hh:31-41).
//
// Layout: workgroup = one 64-lane wavefront = one chain.  The chain's int8
// lattice (column-major, l = i + rows*j) lives in LDS when it fits, else in HBM.
// The MT19937-64 state lives in LDS and is twisted by all 64 lanes together
// (two dependency-free phases of 156 words); every other step is a scalar walk
// that all lanes execute in lockstep with scalar (readfirstlane) operands.
//
// Exactness: a site is drawn by Lemire's method on one engine word; the
// Metropolis decision for class k (old spin, neighbour sum) is either "dE < 0,
// accept without a draw" or "draw one word x and accept iff x < thr[k]", where
// thr[k] is the first word whose libstdc++ canonical double is >= exp(-dE*beta)
// (computed on the host with the same std::exp), so no exp and no
// int->double conversion happens per attempt.  The FP64 tails (running energy
// and the row/column formation-energy fold) use explicitly rounded operations.
#include "internal.hh"

namespace mcg {

__device__ __forceinline__ uint64_t rfl64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int32_t rfl32(int32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// Twist of the 312-word state by one wavefront.  Phase 1 (k < 156) reads only
// old words; phase 2 (k >= 156) reads old x[k], x[k+1] and new x[k-156], x[0].
// Within a phase all reads complete before any write.
__device__ void twist_wave(uint64_t *mt, int lane) {
  const int a0 = lane, a1 = lane + 64, a2 = lane + 128;
  uint64_t v0 = mt_twist_word(mt[a0], mt[a0 + 1], mt[a0 + MT_M]);
  uint64_t v1 = mt_twist_word(mt[a1], mt[a1 + 1], mt[a1 + MT_M]);
  uint64_t v2 = 0;
  if (a2 < MT_N - MT_M) v2 = mt_twist_word(mt[a2], mt[a2 + 1], mt[a2 + MT_M]);
  __syncthreads();
  mt[a0] = v0;
  mt[a1] = v1;
  if (a2 < MT_N - MT_M) mt[a2] = v2;
  __syncthreads();
  const int b0 = 156 + lane, b1 = 220 + lane, b2 = 284 + lane;
  v0 = mt_twist_word(mt[b0], mt[b0 + 1], mt[b0 - MT_M]);
  v1 = mt_twist_word(mt[b1], mt[b1 + 1], mt[b1 - MT_M]);
  if (b2 < MT_N) v2 = mt_twist_word(mt[b2], mt[b2 == MT_N - 1 ? 0 : b2 + 1], mt[b2 - MT_M]);
  __syncthreads();
  mt[b0] = v0;
  mt[b1] = v1;
  if (b2 < MT_N) mt[b2] = v2;
  __syncthreads();
}

template <bool LDS_LATTICE>
__global__ __launch_bounds__(64) void exact_kernel(const ExactParams *__restrict__ Pp,
                                                   ExactState *__restrict__ Sp,
                                                   int8_t *__restrict__ lat_pool,
                                                   mcg_sample *__restrict__ series,
                                                   int64_t series_cap,
                                                   const uint8_t *__restrict__ active,
                                                   int64_t n_passes, int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  const int lane = threadIdx.x;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint64_t mt[MT_N];

  const ExactParams *P = Pp + c;
  ExactState *St = Sp + c;
  const int32_t rows = P->rows, cols = P->cols, n = P->n;
  const uint64_t magic = P->div_magic;
  const uint32_t neg_mask = P->neg_mask, full_mask = P->full_mask;
  const int fold = P->fold;
  const int self_nb = P->self_nb;
  int8_t *g_lat = lat_pool + P->lat_off;
  const int lat_bytes_al = (n + 15) & ~15;
  int8_t *lat = LDS_LATTICE ? reinterpret_cast<int8_t *>(smem) : g_lat;
  int32_t *scratch = reinterpret_cast<int32_t *>(smem + (LDS_LATTICE ? lat_bytes_al : 0));

  for (int i = lane; i < MT_N; i += 64) mt[i] = St->mt[i];
  if (LDS_LATTICE)
    for (int i = lane; i < n; i += 64) lat[i] = g_lat[i];
  __syncthreads();

  int pos = (int)St->pos;
  int64_t n_pass = St->n_pass, n_acc = St->n_accept, n_rej = St->n_reject;
  int64_t S = St->S, M = St->M, n_samples = St->n_samples;
  double e_run = St->e_run;
  const uint64_t N64 = (uint64_t)n;

  auto next_word = [&]() -> uint64_t {
    if (pos >= MT_N) {
      twist_wave(mt, lane);
      pos = 0;
    }
    uint64_t w = rfl64(mt[pos]);
    ++pos;
    return mt_temper(w);
  };

  for (int64_t p = 0; p < n_passes; ++p) {
    for (int32_t step = 0; step < n; ++step) {
      // propose: l = uniform_int_distribution<long>(0, n-1) (Lemire)
      uint64_t x = next_word();
      uint64_t lo = x * N64;
      uint64_t hi = __umul64hi(x, N64);
      if (lo < N64) {
        const uint64_t t = (0ULL - N64) % N64;
        while (lo < t) {
          x = next_word();
          lo = x * N64;
          hi = __umul64hi(x, N64);
        }
      }
      const uint32_t l = (uint32_t)hi;
      const uint32_t j = (uint32_t)(((uint64_t)l * magic) >> 40);
      const uint32_t i = l - j * (uint32_t)rows;
      const uint32_t iu = (i + 1 == (uint32_t)rows) ? 0u : i + 1;
      const uint32_t id = (i == 0) ? (uint32_t)rows - 1 : i - 1;
      const uint32_t ju = (j + 1 == (uint32_t)cols) ? 0u : j + 1;
      const uint32_t jd = (j == 0) ? (uint32_t)cols - 1 : j - 1;
      const int s = rfl32(lat[l]);
      const int nsum = rfl32(lat[iu + rows * j] + lat[id + rows * j] + lat[i + rows * ju] +
                             lat[i + rows * jd]);
      const int k = (s > 0 ? 5 : 0) + ((nsum + 4) >> 1);
      bool acc;
      if ((neg_mask >> k) & 1u) {
        acc = true;
      } else {
        const uint64_t x2 = next_word();
        acc = ((full_mask >> k) & 1u) || (x2 < P->thr[k]);
      }
      if (acc) {
        if (lane == 0) lat[l] = (int8_t)(-s);
        if (!LDS_LATTICE) __threadfence();
        // the bond sum's self terms (a 1-wide lattice is its own neighbour) never
        // change; the reference's dE keeps them (model.hh:312-314), S does not
        S += -2LL * s * (nsum - self_nb * s);
        M += -2LL * s;
        e_run = __dadd_rn(e_run, P->dq[k]);
        ++n_acc;
      } else {
        ++n_rej;
      }
    }
    ++n_pass;
    if (n_pass % sample_period == 0) {
      double ef = 0.0;
      if (fold) {
        __syncthreads();
        for (int r = lane; r < rows; r += 64) {
          const int ru = (r + 1 == rows) ? 0 : r + 1;
          int d = 0;
          for (int q = 0; q < cols; ++q) d += lat[r + rows * q] * lat[ru + rows * q];
          scratch[r] = d;
        }
        for (int q = lane; q < cols; q += 64) {
          const int qu = (q + 1 == cols) ? 0 : q + 1;
          int d = 0;
          for (int r = 0; r < rows; ++r) d += lat[r + rows * q] * lat[r + rows * qu];
          scratch[rows + q] = d;
        }
        __syncthreads();
        if (lane == 0) {
          // model.hh:276-285: sequential fold of fl(-J * dot) over rows, then columns
          double e = 0.0;
          for (int r = 0; r < rows + cols; ++r)
            e = __dadd_rn(e, __dmul_rn(P->negJ, (double)scratch[r]));
          ef = e;
        }
        __syncthreads();
      }
      if (lane == 0 && n_samples < series_cap) {
        mcg_sample rec;
        rec.pass = n_pass;
        rec.S1 = S;
        rec.S2 = 0;
        rec.M = M;
        rec.e_fold = ef;
        rec.e_run = e_run;
        series[(int64_t)c * series_cap + n_samples] = rec;
      }
      ++n_samples;
    }
  }

  __syncthreads();
  for (int i = lane; i < MT_N; i += 64) St->mt[i] = mt[i];
  if (LDS_LATTICE)
    for (int i = lane; i < n; i += 64) g_lat[i] = lat[i];
  if (lane == 0) {
    St->pos = pos;
    St->n_pass = n_pass;
    St->n_accept = n_acc;
    St->n_reject = n_rej;
    St->S = S;
    St->M = M;
    St->e_run = e_run;
    St->n_samples = n_samples;
  }
}

void launch_exact(const ExactParams *P, ExactState *S, int8_t *lat_pool, mcg_sample *series,
                  int64_t series_cap, const uint8_t *active, int n_chains, int64_t n_passes,
                  int64_t sample_period, int lds_bytes, bool lattice_in_lds,
                  hipStream_t stream) {
  if (lattice_in_lds) {
    hipLaunchKernelGGL(exact_kernel<true>, dim3(n_chains), dim3(64), lds_bytes, stream, P, S,
                       lat_pool, series, series_cap, active, n_passes, sample_period);
  } else {
    hipLaunchKernelGGL(exact_kernel<false>, dim3(n_chains), dim3(64), lds_bytes, stream, P, S,
                       lat_pool, series, series_cap, active, n_passes, sample_period);
  }
}

void exact_set_max_lds(int bytes) {
  hipFuncSetAttribute(reinterpret_cast<const void *>(exact_kernel<true>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  hipFuncSetAttribute(reinterpret_cast<const void *>(exact_kernel<false>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
