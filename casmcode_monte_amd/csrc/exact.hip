// exact.hip -- EXACT mode: the reference's sequential random-site Metropolis
// chain, reproduced bit for bit on one wavefront per chain.
//
// Reference: the main loop of basic_occupation_metropolis.hh:381-422 driving the
// SGC Ising calculator (basic_semigrand_canonical.hh:308-320 propose/apply,
// :178-192 potential delta, model.hh:305-345 formation-energy delta) with
// metropolis_acceptance (metropolis.hh:25-35) and RandomNumberGenerator over
// std::mt19937_64 (RandomNumberGenerator.hh:31-41).
//
// Layout: one chain per workgroup.  The chain's int8 lattice (column-major,
// l = i + rows*j) lives in LDS when it fits, else in HBM.  The scalar walk
// (exact_kernel) is one 64-lane wavefront that twists the MT19937-64 state in
// LDS itself (two dependency-free phases of 156 words) and executes every other
// step in lockstep with scalar (readfirstlane) operands.  The block walk
// (exact_block_kernel) has two wavefronts: wave 0 walks, wave 1 twists the
// engine's next generation while wave 0 consumes the current one (mt_producer,
// walk.hh).
//
// Exactness: a site is drawn by Lemire's method on one engine word; the
// Metropolis decision for class k (old spin, neighbour sum) is either "dE < 0,
// accept without a draw" or "draw one word x and accept iff x < thr[k]", where
// thr[k] is the first word whose libstdc++ canonical double is >= exp(-dE*beta)
// (computed on the host with the same std::exp), so no exp and no
// int->double conversion happens per attempt.  The FP64 tails (running energy
// and the row/column formation-energy fold) use explicitly rounded operations.
#include "internal.hh"
#include "walk.hh"

namespace mcg {

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
                                                   int64_t n_passes, int64_t n_steps,
                                                   int64_t sample_period) {
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

  // attempts of this launch: n_passes whole passes and n_steps more, from the
  // chain's position `step` in its current pass
  int32_t step = (int32_t)St->step;
  const int64_t total = n_passes * (int64_t)n + n_steps;
  for (int64_t done = 0; done < total; ++done) {
    {
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
    if (++step < n) continue;
    step = 0;
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
        mcg_sample rec = {};
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
    St->step = step;
    St->n_pass = n_pass;
    St->n_accept = n_acc;
    St->n_reject = n_rej;
    St->S = S;
    St->M = M;
    St->e_run = e_run;
    St->n_samples = n_samples;
  }
}

// ---------------------------------------------------------------------------
// Block-speculative walk (LDS lattice).
//
// The stream is consumed in blocks.  For a block starting at stream position
// P, lane t evaluates position P+t as if it were a site draw: its site (Lemire
// on word x[P+t]), the old spin and neighbour sum from the block-start lattice,
// the class k, whether the attempt would draw a uniform (dE >= 0) and whether
// it would accept (x[P+t+1] < thr[k]).  The real sequence follows in closed
// form: position t is an attempt; the next attempt is at t+1, or t+2 when t
// drew a uniform.  Every attempt lane then marks its site in the lattice byte
// (a draw count in bits 1..7; positions that are uniform draws mark nothing),
// so an attempt whose own site was drawn twice or whose neighbourhood holds
// another attempt's site is flagged as a potential conflict: the block stops
// before such a lane (it will be re-evaluated as lane 0 of the next block, on
// the updated lattice).  Hence
// every attempt the pass accepts saw exactly the lattice the sequential chain
// would see.  Flips are committed at the end of the block; the integer
// observables are block reductions; the FP64 running energy is added in attempt
// order.  A lane 0 whose Lemire product needs the rejection test takes the
// one-attempt scalar path.
// ---------------------------------------------------------------------------
constexpr uint32_t SPIN_BIT = 1u, MARK_A = 2u;

__device__ __forceinline__ int spin_of(uint32_t b) { return (b & SPIN_BIT) ? 1 : -1; }

__global__ __launch_bounds__(128) void exact_block_kernel(const ExactParams *__restrict__ Pp,
                                                         ExactState *__restrict__ Sp,
                                                         int8_t *__restrict__ lat_pool,
                                                         mcg_sample *__restrict__ series,
                                                         int64_t series_cap,
                                                         const uint8_t *__restrict__ active,
                                                         int64_t n_passes, int64_t n_steps,
                                                         int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  const int lane = threadIdx.x & 63;
  const bool producer = threadIdx.x >= 64;  // wave 1 twists the engine; wave 0 walks
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint64_t raw[3][MT_N];
  __shared__ uint64_t ring[3][MT_N];
  __shared__ uint64_t thr_s[EXACT_NCLASS];
  __shared__ double dq_s[EXACT_NCLASS];
  __shared__ int done;

  const ExactParams *P = Pp + c;
  ExactState *St = Sp + c;
  const int32_t rows = P->rows, cols = P->cols, n = P->n;
  const uint64_t magic = P->div_magic;
  const uint32_t neg_mask = P->neg_mask, full_mask = P->full_mask;
  const int fold = P->fold, self_nb = P->self_nb, W = P->window;
  int8_t *g_lat = lat_pool + P->lat_off;
  const int lat_bytes_al = (n + 15) & ~15;
  uint8_t *lat = reinterpret_cast<uint8_t *>(smem);
  uint32_t *latw = reinterpret_cast<uint32_t *>(smem);
  int32_t *scratch = reinterpret_cast<int32_t *>(smem + lat_bytes_al);

  if (!producer) {
    for (int i = lane; i < MT_N; i += 64) {
      raw[0][i] = St->mt[i];
      ring[0][i] = mt_temper(raw[0][i]);
    }
    if (lane < EXACT_NCLASS) {
      thr_s[lane] = P->thr[lane];
      dq_s[lane] = P->dq[lane];
    }
    for (int i = lane; i < lat_bytes_al; i += 64) lat[i] = (i < n && g_lat[i] > 0) ? 1u : 0u;
    if (lane == 0) done = 0;
  }
  __syncthreads();
  if (producer) {
    mt_producer(raw, ring, &done, lane);
    return;
  }

  const int64_t pos0 = St->pos;
  int64_t pos = pos0;  // absolute stream position; generation g holds [312g, 312g+312)
  MtRing3 eng;         // engine generations from the producer wave
  int64_t n_pass = St->n_pass, n_acc = St->n_accept, n_rej = St->n_reject;
  int64_t S = St->S, M = St->M, n_samples = St->n_samples;
  double e_run = St->e_run;
  const uint64_t N64 = (uint64_t)n;

  auto ensure = [&](int64_t upto) { eng.ensure(upto); };
  const uint64_t *ringf = &ring[0][0];
  auto ring_index = [&](int64_t q) -> int { return eng.index(q); };

  // attempts of this launch: n_passes whole passes and n_steps more, from the
  // chain's position `step` in its current pass
  int32_t step = (int32_t)St->step;
  int64_t remaining = n_passes * (int64_t)n + n_steps;
  while (remaining > 0) {
    ensure(pos + W + 1);
    const bool inwin = lane < W;
    const int64_t q = pos + lane;
    const uint64_t x = ringf[ring_index(inwin ? q : pos)];
    const uint64_t x1 = ringf[ring_index(inwin ? q + 1 : pos)];
    // x * n (n < 2^24) from two 32x32->64 products: site = high word, lo = low
    const uint64_t p0 = (uint64_t)(uint32_t)x * (uint32_t)n;
    const uint64_t p1 = (uint64_t)(uint32_t)(x >> 32) * (uint32_t)n + (p0 >> 32);
    const uint64_t lo = (p1 << 32) | (uint32_t)p0;
    const uint32_t site = (uint32_t)(p1 >> 32);
    const bool special = inwin && (lo < N64);
    const bool valid = inwin && !special;
    const uint32_t j = (uint32_t)(((uint64_t)site * magic) >> 40);
    const uint32_t i = site - j * (uint32_t)rows;
    const uint32_t iu = (i + 1 == (uint32_t)rows) ? 0u : i + 1;
    const uint32_t id = (i == 0) ? (uint32_t)rows - 1 : i - 1;
    const uint32_t ju = (j + 1 == (uint32_t)cols) ? 0u : j + 1;
    const uint32_t jd = (j == 0) ? (uint32_t)cols - 1 : j - 1;
    const uint32_t sh = (site & 3u) * 8u;
    const uint32_t a0 = site, a1 = iu + rows * j, a2 = id + rows * j;
    const uint32_t a3 = i + rows * ju, a4 = i + rows * jd;
    // at a block's start every mark is clear, so each byte is its spin bit:
    // the class (s > 0) * 5 + (nsum + 4) / 2 is 5 b0 + b1 + b2 + b3 + b4
    const int b0 = lat[a0], nup = lat[a1] + lat[a2] + lat[a3] + lat[a4];
    const int s = 2 * b0 - 1, nsum = 2 * nup - 4;
    const int k = 5 * b0 + nup;
    const bool nd = !((neg_mask >> k) & 1u);
    const uint64_t ND = __ballot(valid && nd);
    const uint64_t SP = __ballot(special);

    // Attempt positions of the window in closed form: position t is a uniform
    // draw iff t-1 is an attempt that draws (ND), so inside every run of ND bits
    // the attempts sit at even offsets from the run start.  E = attempts that
    // draw, U = E << 1 = uniform positions, attempts = ~U.
    const uint64_t ALT = 0x5555555555555555ull;
    const uint64_t run_start = ND & ~(ND << 1);
    const uint64_t odd_runs = ((ND + (run_start & ~ALT)) ^ ND) & ND;
    const uint64_t E = ND & (ALT ^ odd_runs);
    const uint64_t A = ~(E << 1);
    // count the site draws of the attempt positions in bits 1..7 of each site's
    // byte (at most 64 draws); the wave's LDS operations execute in issue order,
    // so the reads below see every lane's increment (the asm statement keeps
    // the compiler from moving them above the atomics)
    const bool att = valid && ((A >> lane) & 1ull);
    if (att) atomicAdd(&latw[site >> 2], MARK_A << sh);
    asm volatile("" ::: "memory");
    // the class tables go out with the mark reads: one round trip (the asm
    // takes all seven values, so none of the reads moves into a branch)
    uint32_t m0 = lat[a0], m1 = lat[a1], m2 = lat[a2], m3 = lat[a3], m4 = lat[a4];
    uint64_t thr_k = thr_s[k];
    double dqv = dq_s[k];
    asm volatile("" : "+v"(m0), "+v"(m1), "+v"(m2), "+v"(m3), "+v"(m4), "+v"(thr_k), "+v"(dqv));
    const bool acc = !nd || ((full_mask >> k) & 1u) || (x1 < thr_k);
    const uint64_t AC = __ballot(valid && acc);
    // potential conflict: own site drawn twice, or a neighbour drawn at all
    const bool pc = att && ((m0 & 0xFCu) || ((m1 | m2 | m3 | m4) & 0xFEu));
    const uint64_t STOP = __ballot(pc || special || !inwin);
    // execute attempts up to (not including) the first flagged attempt t > 0
    const uint64_t stops = A & STOP & ~1ull;
    uint64_t exec = (SP & 1ull) ? 0ull : (A & (stops ? ((1ull << __builtin_ctzll(stops)) - 1) : ~0ull));
    int attempts = __popcll(exec);
    // the block ends at the end of the pass or of the launch's attempts
    const int lim = (int)(remaining < (int64_t)(n - step) ? remaining : (int64_t)(n - step));
    if (attempts > lim) {
      uint64_t rest = exec;
      for (int r = 0; r < lim; ++r) rest &= rest - 1;
      exec ^= rest;
      attempts = lim;
    }
    int t = 0;
    if (exec) {
      const int last = 63 - __builtin_clzll(exec);
      t = last + 1 + (int)((ND >> last) & 1ull);
    }
    const uint64_t F = exec & AC;

    // commit the block: in-order FP64 running energy and the integer deltas of
    // the flipped lanes (packed: dS + 16 in bits 0..15, dM + 2 in bits 16..31)
    const bool flipped = (F >> lane) & 1ull;
    const uint32_t packed = (uint32_t)(-2 * s * (nsum - self_nb * s) + 16) |
                            ((uint32_t)(-2 * s + 2) << 16);
    uint32_t psum = 0;
    for (uint64_t f = F; f; f &= f - 1) {
      const int b = __builtin_ctzll(f);
      e_run = __dadd_rn(e_run, readlane_f64(dqv, b));
      psum += __builtin_amdgcn_readlane(packed, b);
    }
    const int nflip = __popcll(F);
    S += (int64_t)(psum & 0xFFFFu) - 16 * nflip;
    M += (int64_t)(psum >> 16) - 2 * nflip;
    n_acc += nflip;
    n_rej += attempts - nflip;
    // one add per attempt lane takes its own mark back out of its site's byte
    // and, for a flipped site (drawn once in the window, so its byte is
    // b0 + MARK_A), turns the spin bit: the byte never goes below zero, so
    // no borrow reaches the next site
    if (att) {
      const uint32_t dv = flipped ? (b0 ? 0u - (MARK_A + SPIN_BIT) : SPIN_BIT - MARK_A) : 0u - MARK_A;
      atomicAdd(&latw[site >> 2], dv << sh);
    }
    asm volatile("" ::: "memory");
    pos += t;
    step += attempts;
    remaining -= attempts;

    // lane 0 special: exec was empty, so this is the block's only attempt and
    // `remaining` was > 0 at the top of the loop
    if (SP & 1ull) {
      // one attempt with the full Lemire rejection loop (scalar walk)
      uint64_t xs = ringf[ring_index(pos)];
      ++pos;
      uint64_t los = xs * N64, his = __umul64hi(xs, N64);
      if (los < N64) {
        const uint64_t th = (0ULL - N64) % N64;
        while (los < th) {
          ensure(pos + 1);
          xs = ringf[ring_index(pos)];
          ++pos;
          los = xs * N64;
          his = __umul64hi(xs, N64);
        }
      }
      const uint32_t l = (uint32_t)his;
      const uint32_t jj = (uint32_t)(((uint64_t)l * magic) >> 40);
      const uint32_t ii = l - jj * (uint32_t)rows;
      const uint32_t iiu = (ii + 1 == (uint32_t)rows) ? 0u : ii + 1;
      const uint32_t iid = (ii == 0) ? (uint32_t)rows - 1 : ii - 1;
      const uint32_t jju = (jj + 1 == (uint32_t)cols) ? 0u : jj + 1;
      const uint32_t jjd = (jj == 0) ? (uint32_t)cols - 1 : jj - 1;
      const int ss = spin_of(lat[l]);
      const int ns = spin_of(lat[iiu + rows * jj]) + spin_of(lat[iid + rows * jj]) +
                     spin_of(lat[ii + rows * jju]) + spin_of(lat[ii + rows * jjd]);
      const int kk = (ss > 0 ? 5 : 0) + ((ns + 4) >> 1);
      bool a;
      if ((neg_mask >> kk) & 1u) {
        a = true;
      } else {
        ensure(pos + 1);
        const uint64_t x2 = ringf[ring_index(pos)];
        ++pos;
        a = ((full_mask >> kk) & 1u) || (x2 < thr_s[kk]);
      }
      if (a) {
        if (lane == 0) atomicXor(&latw[l >> 2], SPIN_BIT << ((l & 3u) * 8u));
        S += -2LL * ss * (ns - self_nb * ss);
        M += -2LL * ss;
        e_run = __dadd_rn(e_run, dq_s[kk]);
        ++n_acc;
      } else {
        ++n_rej;
      }
      // one wavefront walks: its LDS operations execute in issue order, and
      // the asm keeps the compiler from reordering them (no workgroup barrier
      // here: the producer wave takes part only in the engine hand-offs)
      asm volatile("" ::: "memory");
      ++step;
      --remaining;
    }

    if (step == n) {
      step = 0;
      ++n_pass;
      if (n_pass % sample_period == 0) {
        double ef = 0.0;
        if (fold) {
          for (int r = lane; r < rows; r += 64) {
            const int ru = (r + 1 == rows) ? 0 : r + 1;
            int d = 0;
            for (int qq = 0; qq < cols; ++qq)
              d += spin_of(lat[r + rows * qq]) * spin_of(lat[ru + rows * qq]);
            scratch[r] = d;
          }
          for (int qq = lane; qq < cols; qq += 64) {
            const int qu = (qq + 1 == cols) ? 0 : qq + 1;
            int d = 0;
            for (int r = 0; r < rows; ++r)
              d += spin_of(lat[r + rows * qq]) * spin_of(lat[r + rows * qu]);
            scratch[rows + qq] = d;
          }
          asm volatile("" ::: "memory");
          if (lane == 0) {
            double e = 0.0;  // model.hh:276-285
            for (int r = 0; r < rows + cols; ++r)
              e = __dadd_rn(e, __dmul_rn(P->negJ, (double)scratch[r]));
            ef = e;
          }
          asm volatile("" ::: "memory");
        }
        if (lane == 0 && n_samples < series_cap) {
          mcg_sample rec = {};
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
  }

  // release the producer: it is waiting at its next barrier, which this one
  // matches; it then reads the barrier's number and returns
  eng.finish(&done, lane);
  // write back in the std::mt19937_64 layout: the generation holding the last
  // consumed output, with index in (0, 312]
  int64_t g_fin = 0, p_fin = pos0;
  if (pos != pos0) {
    g_fin = (pos - 1) / MT_N;
    p_fin = pos - (int64_t)MT_N * g_fin;
  }
  for (int i2 = lane; i2 < MT_N; i2 += 64) St->mt[i2] = raw[g_fin % 3][i2];
  for (int i2 = lane; i2 < n; i2 += 64) g_lat[i2] = (lat[i2] & SPIN_BIT) ? (int8_t)1 : (int8_t)-1;
  if (lane == 0) {
    St->pos = p_fin;
    St->step = step;
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
                  int64_t n_steps, int64_t sample_period, int lds_bytes, bool lattice_in_lds,
                  bool block_walk, hipStream_t stream) {
  if (lattice_in_lds && block_walk) {
    hipLaunchKernelGGL(exact_block_kernel, dim3(n_chains), dim3(128), lds_bytes, stream, P, S,
                       lat_pool, series, series_cap, active, n_passes, n_steps, sample_period);
  } else if (lattice_in_lds) {
    hipLaunchKernelGGL(exact_kernel<true>, dim3(n_chains), dim3(64), lds_bytes, stream, P, S,
                       lat_pool, series, series_cap, active, n_passes, n_steps, sample_period);
  } else {
    hipLaunchKernelGGL(exact_kernel<false>, dim3(n_chains), dim3(64), lds_bytes, stream, P, S,
                       lat_pool, series, series_cap, active, n_passes, n_steps, sample_period);
  }
}

// The block walk's static LDS (three raw and three tempered engine
// generations, ~15 KB) leaves it 12 KB less dynamic LDS than the scalar walk.
hipError_t exact_set_max_lds(int bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(exact_block_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes - 12 * 1024);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void *>(exact_kernel<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(exact_kernel<false>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
