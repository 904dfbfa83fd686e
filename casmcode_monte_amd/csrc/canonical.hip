// canonical.hip -- CANONICAL mode: the reference's canonical pair-swap chain
// (Kawasaki exchanges between any two sites of different species), reproduced
// bit for bit on one wavefront per chain; a second wavefront of the workgroup
// twists the engine's next generation meanwhile (mt_producer, walk.hh).
//
// Reference: the generic loop of occupation_metropolis.hh:87-153 with
// propose_canonical_event (OccEventProposal.hh:109-236) over an OccLocation
// (OccLocation.hh:254-259 choose_mol; OccLocation.cc:216-283 apply) and
// metropolis_acceptance (metropolis.hh:25-35), on a binary lattice with one
// asymmetric unit: species 0 = spin -1, species 1 = spin +1, one canonical
// swap type ((0,0),(0,1)).  The potential is the square-lattice Ising
// formation energy with the multi-site delta of model.hh:354-379.
//
// Per step the reference draws: one word for the swap-type choice (with a
// single swap type it always picks it), one Lemire word for the position in
// the species-0 list (site a), one for the species-1 list (site b), and one
// more only when dE >= 0.  dE depends on (Sa, Sb'): the neighbour sum of a
// before the swap and of b after a has flipped, so the acceptance is a per-class
// word threshold as in the EXACT mode.
//
// Device OccLocation: the two candidate lists as 16-bit site indices in LDS
// (species 0 at [0, nA), species 1 at [nA, n)); Mol ids equal site indices
// because every site mutates.  An accepted swap (a at list-0 slot ia, b at
// list-1 slot ib) leaves list0[ia] = old list0[nA-1], list0[nA-1] = b and
// list1[ib] = a, which is what the reference's remove-by-swap-with-back then
// push_back sequence gives.
//
// Block-speculative walk: lane t evaluates stream position P+t as if a step
// started there (its two list slots, sites, classes and acceptance from the
// block-start state).  Steps take 3 or 4 words, so the real step starts are
// parsed from the "draws a uniform" ballot.  A step is exact if no earlier
// accepted step of the block touched its footprint (its two sites and their
// neighbours) or the list slots it reads: accepted steps write their lane index
// into per-site marks with an LDS atomic minimum, every step reads the marks of
// its footprint, and the block executes the steps before the first one that
// sees an earlier mark (or reads slot nA-1 after an earlier acceptance).
// Because only earlier accepted steps can flag a step, the first flagged step
// is exact in its flag: all steps before it saw the sequential chain's state.
#include "internal.hh"
#include "walk.hh"

namespace mcg {

namespace {

constexpr uint32_t NO_MARK = 0xFFFFFFFFu;

__device__ __forceinline__ int spin_bit(uint8_t v) { return v ? 1 : -1; }

struct Nbrs {
  uint32_t up, dn, rt, lf;
};

__device__ __forceinline__ Nbrs nbrs_of(uint32_t l, int32_t rows, int32_t cols, uint64_t magic) {
  const uint32_t j = (uint32_t)(((uint64_t)l * magic) >> 40);
  const uint32_t i = l - j * (uint32_t)rows;
  const uint32_t iu = (i + 1 == (uint32_t)rows) ? 0u : i + 1;
  const uint32_t id = (i == 0) ? (uint32_t)rows - 1 : i - 1;
  const uint32_t ju = (j + 1 == (uint32_t)cols) ? 0u : j + 1;
  const uint32_t jd = (j == 0) ? (uint32_t)cols - 1 : j - 1;
  return {iu + (uint32_t)rows * j, id + (uint32_t)rows * j, i + (uint32_t)rows * ju,
          i + (uint32_t)rows * jd};
}

}  // namespace

__global__ __launch_bounds__(128) void canonical_kernel(const CanonParams *__restrict__ Pp,
                                                       ExactState *__restrict__ Sp,
                                                       int8_t *__restrict__ lat_pool,
                                                       int32_t *__restrict__ list_pool,
                                                       mcg_sample *__restrict__ series,
                                                       int64_t series_cap,
                                                       const uint8_t *__restrict__ active,
                                                       int64_t n_passes,
                                                       int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  const int lane = threadIdx.x & 63;
  const bool producer = threadIdx.x >= 64;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint64_t raw[3][MT_N];
  __shared__ uint64_t ring[3][MT_N];
  __shared__ int done;
  __shared__ uint64_t thr_s[CANON_NCLASS];
  __shared__ double dq_s[CANON_NCLASS];

  const CanonParams *P = Pp + c;
  ExactState *St = Sp + c;
  const int32_t rows = P->rows, cols = P->cols, n = P->n;
  const uint32_t nA = (uint32_t)P->nA, nB = (uint32_t)P->nB;
  const uint64_t magic = P->div_magic;
  const uint32_t neg_mask = P->neg_mask, full_mask = P->full_mask, mmask = P->mark_mask;
  const int fold = P->fold, self_nb = P->self_nb, W = P->window;
  int8_t *g_lat = lat_pool + P->lat_off;
  int32_t *g_list = list_pool + P->lat_off;
  const int lat_al = (n + 15) & ~15;
  uint8_t *lat = smem;
  uint32_t *marks = reinterpret_cast<uint32_t *>(smem + lat_al);
  uint16_t *list0 = reinterpret_cast<uint16_t *>(smem + lat_al + 4 * (mmask + 1));
  uint16_t *list1 = list0 + nA;
  int32_t *scratch = reinterpret_cast<int32_t *>(smem + lat_al + 4 * (mmask + 1) +
                                                 ((2 * n + 15) & ~15));

  if (!producer) {
    for (int i = lane; i < MT_N; i += 64) {
      raw[0][i] = St->mt[i];
      ring[0][i] = mt_temper(raw[0][i]);
    }
    if (lane < CANON_NCLASS) {
      thr_s[lane] = P->thr[lane];
      dq_s[lane] = P->dq[lane];
    }
    for (int i = lane; i < n; i += 64) {
      lat[i] = g_lat[i] > 0 ? 1u : 0u;
      list0[i] = (uint16_t)g_list[i];
    }
    for (uint32_t i = lane; i <= mmask; i += 64) marks[i] = NO_MARK;
    if (lane == 0) done = 0;
  }
  __syncthreads();
  if (producer) {  // wave 1 twists the engine (walk.hh); wave 0 walks
    mt_producer(raw, ring, &done, lane);
    return;
  }

  const int64_t pos0 = St->pos;
  int64_t pos = pos0;
  MtRing3 eng;  // engine generations from the producer wave
  int64_t n_pass = St->n_pass, n_acc = St->n_accept, n_rej = St->n_reject;
  int64_t S = St->S, M = St->M, n_samples = St->n_samples;
  double e_run = St->e_run;
  const uint64_t NA64 = nA, NB64 = nB;

  auto ensure = [&](int64_t upto) { eng.ensure(upto); };
  const uint64_t *ringf = &ring[0][0];
  auto ring_index = [&](int64_t q) -> int { return eng.index(q); };
  // libstdc++ uniform_int_distribution over [0, m-1] with the rejection loop
  auto lemire_scalar = [&](uint64_t m) -> uint32_t {
    ensure(pos + 1);
    uint64_t x = ringf[ring_index(pos)];
    ++pos;
    uint64_t lo = x * m, hi = __umul64hi(x, m);
    if (lo < m) {
      const uint64_t th = (0ULL - m) % m;
      while (lo < th) {
        ensure(pos + 1);
        x = ringf[ring_index(pos)];
        ++pos;
        lo = x * m;
        hi = __umul64hi(x, m);
      }
    }
    return (uint32_t)hi;
  };

  int32_t step = 0;
  int64_t passes_done = 0;
  while (passes_done < n_passes) {
    ensure(pos + W + 3);
    const bool inwin = lane < W;
    const int64_t q = inwin ? pos + lane : pos;
    const uint64_t x1 = ringf[ring_index(q + 1)];
    const uint64_t x2 = ringf[ring_index(q + 2)];
    const uint64_t x3 = ringf[ring_index(q + 3)];
    const uint32_t ia = (uint32_t)__umul64hi(x1, NA64);
    const uint32_t ib = (uint32_t)__umul64hi(x2, NB64);
    const bool special = inwin && ((x1 * NA64) < NA64 || (x2 * NB64) < NB64);
    const bool valid = inwin && !special;
    const uint32_t back = list0[nA - 1];  // block-start list0[nA-1]
    const uint32_t a = list0[ia], b = list1[ib];
    const Nbrs na = nbrs_of(a, rows, cols, magic), nb = nbrs_of(b, rows, cols, magic);
    const int Sa = spin_bit(lat[na.up]) + spin_bit(lat[na.dn]) + spin_bit(lat[na.rt]) +
                   spin_bit(lat[na.lf]);
    const int mult = (nb.up == a) + (nb.dn == a) + (nb.rt == a) + (nb.lf == a);
    const int Sb = spin_bit(lat[nb.up]) + spin_bit(lat[nb.dn]) + spin_bit(lat[nb.rt]) +
                   spin_bit(lat[nb.lf]) + 2 * mult;
    const int k = ((Sa + 4) >> 1) * 5 + ((Sb + 4) >> 1);
    const bool nd = !((neg_mask >> k) & 1u);
    const bool acc = !nd || ((full_mask >> k) & 1u) || (x3 < thr_s[k]);
    const uint64_t ND = __ballot(valid && nd);
    const uint64_t AC = __ballot(valid && acc);
    const uint64_t SP = __ballot(special);

    // step starts: 0, then +3 or +4 (a uniform drawn)
    uint64_t A = 0;
    for (int t = 0; t < W; t += 3 + (int)((ND >> t) & 1ull)) A |= 1ull << t;
    const bool att = valid && ((A >> lane) & 1ull);
    const uint64_t ACC = A & AC;
    const bool accl = att && acc;
    if (accl) {
      atomicMin(&marks[a & mmask], (uint32_t)lane);
      atomicMin(&marks[b & mmask], (uint32_t)lane);
    }
    // the wave's LDS operations execute in issue order: the reads below see
    // every lane's marks (the asm statement keeps the compiler from hoisting them)
    asm volatile("" ::: "memory");
    bool conflict = false;
    if (att) {
      uint32_t mk = min(marks[a & mmask], marks[b & mmask]);
      mk = min(mk, min(min(marks[na.up & mmask], marks[na.dn & mmask]),
                       min(marks[na.rt & mmask], marks[na.lf & mmask])));
      mk = min(mk, min(min(marks[nb.up & mmask], marks[nb.dn & mmask]),
                       min(marks[nb.rt & mmask], marks[nb.lf & mmask])));
      const bool earlier_acc = (ACC & ((1ull << lane) - 1ull)) != 0ull;
      conflict = mk < (uint32_t)lane || (ia == nA - 1 && earlier_acc);
    }
    const uint64_t STOP = __ballot(conflict || special || !inwin);
    const uint64_t stops = A & STOP & ~1ull;
    uint64_t exec = (SP & 1ull) ? 0ull
                                : (A & (stops ? ((1ull << __builtin_ctzll(stops)) - 1) : ~0ull));
    int attempts = __popcll(exec);
    const int to_pass_end = n - step;
    if (attempts > to_pass_end) {
      uint64_t rest = exec;
      for (int r = 0; r < to_pass_end; ++r) rest &= rest - 1;
      exec ^= rest;
      attempts = to_pass_end;
    }
    int used = 0;
    if (exec) {
      const int last = 63 - __builtin_clzll(exec);
      used = last + 3 + (int)((ND >> last) & 1ull);
    }
    const uint64_t F = exec & AC;

    // commit: FP64 running energy in step order, bond-sum deltas
    // (2 (Sa - Sb) + 4 self_nb per swap), occupations and lists
    const bool flipped = (F >> lane) & 1ull;
    const int dS = 2 * (Sa - Sb) + 4 * self_nb;
    int dsum = 0;
    for (uint64_t f = F; f; f &= f - 1) {
      const int bl = __builtin_ctzll(f);
      e_run = __dadd_rn(e_run, readlane_f64(dq_s[k], bl));
      dsum += __builtin_amdgcn_readlane(dS, bl);
    }
    const int nflip = __popcll(F);
    S += dsum;
    n_acc += nflip;
    n_rej += attempts - nflip;
    const uint64_t lower = F & ((1ull << lane) - 1ull);
    const int prev = lower ? 63 - __builtin_clzll(lower) : lane;
    const uint32_t b_prev = (uint32_t)__shfl((int)b, prev, 64);
    if (flipped) {
      lat[a] = 1u;
      lat[b] = 0u;
      list0[ia] = (uint16_t)(lower ? b_prev : back);
      list1[ib] = (uint16_t)a;
    }
    asm volatile("" ::: "memory");
    if (F && lane == 63 - __builtin_clzll(F)) list0[nA - 1] = (uint16_t)b;
    if (accl) {
      marks[a & mmask] = NO_MARK;
      marks[b & mmask] = NO_MARK;
    }
    asm volatile("" ::: "memory");
    pos += used;
    step += attempts;

    if (SP & 1ull) {
      // one step with the full Lemire rejection loops
      ensure(pos + 1);
      ++pos;  // swap-type choice
      const uint32_t sa = lemire_scalar(NA64);
      const uint32_t sb = lemire_scalar(NB64);
      const uint32_t la = list0[sa], lb = list1[sb];
      const Nbrs ma = nbrs_of(la, rows, cols, magic), mb = nbrs_of(lb, rows, cols, magic);
      const int Sa1 = spin_bit(lat[ma.up]) + spin_bit(lat[ma.dn]) + spin_bit(lat[ma.rt]) +
                      spin_bit(lat[ma.lf]);
      const int m1 = (mb.up == la) + (mb.dn == la) + (mb.rt == la) + (mb.lf == la);
      const int Sb1 = spin_bit(lat[mb.up]) + spin_bit(lat[mb.dn]) + spin_bit(lat[mb.rt]) +
                      spin_bit(lat[mb.lf]) + 2 * m1;
      const int kk = ((Sa1 + 4) >> 1) * 5 + ((Sb1 + 4) >> 1);
      bool ok;
      if ((neg_mask >> kk) & 1u) {
        ok = true;
      } else {
        ensure(pos + 1);
        const uint64_t xr = ringf[ring_index(pos)];
        ++pos;
        ok = ((full_mask >> kk) & 1u) || (xr < thr_s[kk]);
      }
      // one wavefront walks: its LDS operations execute in issue order and the
      // asm keeps the compiler from reordering them (the producer wave takes
      // part only in the engine hand-offs)
      asm volatile("" ::: "memory");
      if (ok) {
        if (lane == 0) {
          const uint16_t bk = list0[nA - 1];
          lat[la] = 1u;
          lat[lb] = 0u;
          list0[sa] = bk;
          list0[nA - 1] = (uint16_t)lb;
          list1[sb] = (uint16_t)la;
        }
        S += 2 * (Sa1 - Sb1) + 4 * self_nb;
        e_run = __dadd_rn(e_run, dq_s[kk]);
        ++n_acc;
      } else {
        ++n_rej;
      }
      asm volatile("" ::: "memory");
      ++step;
    }

    if (step == n) {
      step = 0;
      ++passes_done;
      ++n_pass;
      if (n_pass % sample_period == 0) {
        double ef = 0.0;
        if (fold) {
          asm volatile("" ::: "memory");
          for (int r = lane; r < rows; r += 64) {
            const int ru = (r + 1 == rows) ? 0 : r + 1;
            int d = 0;
            for (int qq = 0; qq < cols; ++qq)
              d += spin_bit(lat[r + rows * qq]) * spin_bit(lat[ru + rows * qq]);
            scratch[r] = d;
          }
          for (int qq = lane; qq < cols; qq += 64) {
            const int qu = (qq + 1 == cols) ? 0 : qq + 1;
            int d = 0;
            for (int r = 0; r < rows; ++r)
              d += spin_bit(lat[r + rows * qq]) * spin_bit(lat[r + rows * qu]);
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
          rec.M = M;
          rec.e_fold = ef;
          rec.e_run = e_run;
          series[(int64_t)c * series_cap + n_samples] = rec;
        }
        ++n_samples;
      }
    }
  }

  // release the producer (it waits at the barrier this one matches)
  eng.finish(&done, lane);
  int64_t g_fin = 0, p_fin = pos0;
  if (pos != pos0) {
    g_fin = (pos - 1) / MT_N;
    p_fin = pos - (int64_t)MT_N * g_fin;
  }
  for (int i = lane; i < MT_N; i += 64) St->mt[i] = raw[g_fin % 3][i];
  for (int i = lane; i < n; i += 64) {
    g_lat[i] = lat[i] ? (int8_t)1 : (int8_t)-1;
    g_list[i] = (int32_t)list0[i];
  }
  if (lane == 0) {
    St->pos = p_fin;
    St->n_pass = n_pass;
    St->n_accept = n_acc;
    St->n_reject = n_rej;
    St->S = S;
    St->M = M;
    St->e_run = e_run;
    St->n_samples = n_samples;
  }
}

void launch_canonical(const CanonParams *P, ExactState *S, int8_t *lat_pool, int32_t *list_pool,
                      mcg_sample *series, int64_t series_cap, const uint8_t *active,
                      int n_chains, int64_t n_passes, int64_t sample_period, int lds_bytes,
                      hipStream_t stream) {
  hipLaunchKernelGGL(canonical_kernel, dim3(n_chains), dim3(128), lds_bytes, stream, P, S,
                     lat_pool, list_pool, series, series_cap, active, n_passes, sample_period);
}

hipError_t canonical_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(canonical_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
