// occupant.hip -- OCCUPANT mode: the reference's generic occupant-event
// Metropolis chain over an OccLocation with several asymmetric units, for
// semi-grand canonical single-site events or canonical pair swaps, reproduced
// bit for bit on one wavefront per chain; a second wavefront of the workgroup
// twists the engine's next generation meanwhile (mt_producer, walk.hh).
//
// Reference: the loop of occupation_metropolis.hh:87-153 with
// propose_semigrand_canonical_event (OccEventProposal.hh:261-352) or
// propose_canonical_event (:109-236) over an OccLocation (OccLocation.hh:
// 257-261 choose_mol; OccLocation.cc:39-118 initialize, 216-283 apply) and
// metropolis_acceptance (metropolis.hh:25-35).
//
// Per step the reference draws one word for the swap type (random_real over
// the summed counts of the swap types: n(cand_a) for semi-grand swaps,
// n(cand_a) n(cand_b) for canonical ones), one Lemire word for the slot in
// cand_a's list (and, canonical, one for cand_b's), and one more only when
// dE >= 0.  dE is a function of the swap type and the neighbour sums (Sa, and
// for pairs Sb of b after a changed), so the acceptance is a per-class word
// threshold as in the EXACT mode; the class tables come from the host.
//
// Device OccLocation: every asymmetric unit b with two occupants owns a
// segment [off_b, off_b + size_b) of the LDS list array that holds both of its
// candidate lists: occupant 0's from the front (slot j at off_b + j), occupant
// 1's from the back (slot j at off_b + size_b - 1 - j).  Removing slot i of a
// list moves its last entry into slot i (OccLocation.cc:262-266), and the
// position that frees is exactly where the other list's push_back lands, so a
// transform is two LDS writes.  Entries are site indices; the host converts
// them to and from Mol ids.  Sites of units with one occupant are in no list.
//
// Block walk: lane t evaluates stream position P+t as if a step started there,
// from the block-start state.  Rejected steps change nothing, so every step up
// to and including the first accepted one is exact; the block commits those
// and starts again after them.  (An accepted step changes the list sizes and
// with them every later step's swap choice and Lemire range, so nothing past
// it can be kept.)
#include "internal.hh"
#include "walk.hh"

namespace mcg {

namespace {

// one swap type's lists at the block start: sizes, slot -> position maps
// (position = base + dir * slot) and units.  dir is +1 for a unit's occupant-0
// list (the front of its segment) and -1 for its occupant-1 list, so it is
// also the occupant change of a transformed site (0 -> 1 or 1 -> 0).
struct SwapTab {
  int32_t na, base_a, dir_a, ua;
  int32_t nb, base_b, dir_b, ub;
};

__device__ __forceinline__ int spin_of(uint8_t v) { return (v & 1) ? 1 : -1; }

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

// neighbour spin sums of site l: over all four slots (the reference's delta
// reads a slot that is the site itself with its old spin) and over the slots
// that are other sites (the bond-sum change); `x`, if a neighbour, counts with
// spin sx
__device__ __forceinline__ void nbr_sums(const uint8_t *lat, uint32_t l, const Nbrs &nb,
                                         uint32_t x, int sx, int &all, int &ex) {
  const uint32_t s4[4] = {nb.up, nb.dn, nb.rt, nb.lf};
  all = 0;
  ex = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int s = s4[r] == x ? sx : spin_of(lat[s4[r]]);
    all += s;
    ex += s4[r] == l ? 0 : s;
  }
}

// One proposed event evaluated from a state: the swap type, the chosen sites
// with their list positions and the positions (and entries) of their lists'
// last slots, the class, and the bond-sum and spin-sum changes.
struct Proposal {
  SwapTab e;
  int32_t pa, pba, pb, pbb;
  uint32_t a, la, b, lb;
  int k, dS, dM;
};

// The neighbour sums, class and observable changes of a proposal whose sites
// are known (the site and list-entry reads done).
__device__ __forceinline__ void classify(Proposal &p, int si, int canonical, const uint8_t *lat,
                                         int32_t rows, int32_t cols, uint64_t magic) {
  const int oa = p.e.dir_a < 0 ? 1 : 0;
  const int sa_old = 2 * oa - 1, sa_new = -sa_old;
  int Sa, Sa_ex;
  nbr_sums(lat, p.a, nbrs_of(p.a, rows, cols, magic), 0xFFFFFFFFu, 0, Sa, Sa_ex);
  p.k = si * 25 + ((Sa + 4) >> 1) * 5;
  p.dS = (sa_new - sa_old) * Sa_ex;
  p.dM = sa_new - sa_old;
  if (canonical) {
    const int sb_old = p.e.dir_b < 0 ? 1 : -1, sb_new = -sb_old;
    int Sb, Sb_ex;
    nbr_sums(lat, p.b, nbrs_of(p.b, rows, cols, magic), p.a, sa_new, Sb, Sb_ex);
    p.k += (Sb + 4) >> 1;
    p.dS += (sb_new - sb_old) * Sb_ex;
    p.dM += sb_new - sb_old;
  }
}

// OccLocation.cc:254-283 for a, then for b (the second transform sees the
// first's lists), by one lane: each removal moves its list's last entry into
// the freed slot, and the position the last entry leaves is where the other
// list of the unit appends.  With a and b on one unit, b's list has gained a
// at that position, so b's removal moves a into b's slot and b lands there.
__device__ __forceinline__ void apply_proposal(const Proposal &p, int canonical, uint8_t *lat,
                                               uint16_t *list, int32_t *cnt) {
  list[p.pa] = (uint16_t)p.la;
  list[p.pba] = (uint16_t)p.a;
  lat[p.a] = (uint8_t)(p.e.dir_a > 0 ? 1 : 0);
  atomicAdd(&cnt[p.e.ua], -p.e.dir_a);
  if (canonical) {
    if (p.e.ub == p.e.ua) {
      list[p.pb] = (uint16_t)p.a;
      list[p.pba] = (uint16_t)p.b;
    } else {
      list[p.pb] = (uint16_t)p.lb;
      list[p.pbb] = (uint16_t)p.b;
    }
    lat[p.b] = (uint8_t)(p.e.dir_b > 0 ? 1 : 0);
    atomicAdd(&cnt[p.e.ub], -p.e.dir_b);
  }
}

}  // namespace

__global__ __launch_bounds__(128) void occupant_kernel(const OccParams *__restrict__ Pp,
                                                      const OccTables *__restrict__ Tp,
                                                      ExactState *__restrict__ Sp,
                                                      int8_t *__restrict__ lat_pool,
                                                      int32_t *__restrict__ list_pool,
                                                      mcg_sample *__restrict__ series,
                                                      int64_t series_cap,
                                                      const uint8_t *__restrict__ active,
                                                      int64_t n_passes, int64_t n_steps,
                                                      int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  const int lane = threadIdx.x & 63;
  const bool producer = threadIdx.x >= 64;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint64_t raw[3][MT_N];
  __shared__ uint64_t ring[3][MT_N];
  __shared__ int done;
  __shared__ uint64_t thr_s[OCC_NCLASS];
  __shared__ double dq_s[OCC_NCLASS];
  __shared__ uint8_t flg_s[OCC_NCLASS];
  __shared__ SwapTab tab_s[OCC_MAX_SWAPS];
  __shared__ int32_t cnt_s[OCC_MAX_ASYM];

  const OccParams *P = Pp + c;
  const OccTables *T = Tp + c;
  ExactState *St = Sp + c;
  const int32_t rows = P->rows, cols = P->cols, n = P->n, n_mut = P->n_mut;
  const int canonical = P->canonical, n_swaps = P->n_swaps, W = P->window;
  const uint64_t magic = P->div_magic;
  int8_t *g_lat = lat_pool + P->lat_off;
  int32_t *g_list = list_pool + P->lat_off;
  uint8_t *lat = smem;
  uint16_t *list = reinterpret_cast<uint16_t *>(smem + ((n + 15) & ~15));

  if (!producer) {
    for (int i = lane; i < MT_N; i += 64) {
      raw[0][i] = St->mt[i];
      ring[0][i] = mt_temper(raw[0][i]);
    }
    for (int i = lane; i < OCC_NCLASS; i += 64) {
      thr_s[i] = T->thr[i];
      dq_s[i] = T->dq[i];
      flg_s[i] = T->flags[i];
    }
    if (lane < OCC_MAX_ASYM) cnt_s[lane] = St->occ_cnt0[lane];
    for (int i = lane; i < n; i += 64) lat[i] = (uint8_t)g_lat[i];
    for (int i = lane; i < n_mut; i += 64) list[i] = (uint16_t)g_list[i];
    if (lane == 0) done = 0;
  }
  __syncthreads();
  if (producer) {  // wave 1 twists the engine (walk.hh); wave 0 walks
    mt_producer(raw, ring, &done, lane);
    return;
  }

  // Lane j < n_swaps keeps swap type j's table entry (its two lists' sizes
  // change with every accepted event of their units) and writes it to tab_s
  // when it changes; the cumulative counts tsum[j+1] are exact integers, kept
  // uniform.
  SwapTab mine = {};
  if (lane < n_swaps) {
    const int ca = P->swap_a[lane], cb = P->swap_b[lane];
    const int ua = P->cand_asym[ca], oa = P->cand_occ[ca];
    const int ub = P->cand_asym[cb], ob = P->cand_occ[cb];
    const int32_t c0a = St->occ_cnt0[ua], c0b = St->occ_cnt0[ub];
    const int32_t za = P->asym_size[ua], zb = P->asym_size[ub];
    mine.na = oa ? za - c0a : c0a;
    mine.base_a = P->asym_off[ua] + (oa ? za - 1 : 0);
    mine.dir_a = oa ? -1 : 1;
    mine.ua = ua;
    mine.nb = ob ? zb - c0b : c0b;
    mine.base_b = P->asym_off[ub] + (ob ? zb - 1 : 0);
    mine.dir_b = ob ? -1 : 1;
    mine.ub = ub;
  }

  const int64_t pos0 = St->pos;
  int64_t pos = pos0;
  MtRing3 eng;
  int64_t n_pass = St->n_pass, n_acc = St->n_accept, n_rej = St->n_reject;
  int64_t S = St->S, M = St->M, n_samples = St->n_samples;
  double e_run = St->e_run;
  int64_t step = St->step;
  int64_t frozen = 0;

  auto ensure = [&](int64_t upto) { eng.ensure(upto); };
  const uint64_t *ringf = &ring[0][0];
  auto word = [&](int64_t q) -> uint64_t { return ringf[eng.index(q)]; };
  // libstdc++ uniform_int_distribution over [0, m-1] with the rejection loop
  auto lemire_scalar = [&](uint64_t m) -> uint32_t {
    ensure(pos + 1);
    uint64_t x = word(pos);
    ++pos;
    uint64_t lo = x * m, hi = __umul64hi(x, m);
    if (lo < m) {
      const uint64_t th = (0ULL - m) % m;
      while (lo < th) {
        ensure(pos + 1);
        x = word(pos);
        ++pos;
        lo = x * m;
        hi = __umul64hi(x, m);
      }
    }
    return (uint32_t)hi;
  };

  uint32_t ts[OCC_MAX_SWAPS];
  double Ttot = 0.0;
  // OccEventProposal.hh:127-139 / 276-288: the cumulative counts from the
  // lanes' entries (n(cand_a), or n(cand_a) n(cand_b) for canonical swaps)
  auto publish = [&]() {
    uint32_t w = 0;
    if (lane < n_swaps) {
      tab_s[lane] = mine;
      w = canonical ? (uint32_t)mine.na * (uint32_t)mine.nb : (uint32_t)mine.na;
    }
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < OCC_MAX_SWAPS; ++j) {
      run += __builtin_amdgcn_readlane(w, j);
      ts[j] = run;
    }
    Ttot = (double)run;
    asm volatile("" ::: "memory");
  };
  // the swap type of a draw: the first j with rand < tsum[j+1]; for an integer
  // t, rand < t exactly when floor(rand) < t, so the comparisons are integer
  // ones: the number of tsum[j+1] <= floor(rand) (they do not decrease, and
  // rand < tsum[n_swaps])
  auto choose = [&](double rand) -> int {
    const uint32_t r = (uint32_t)rand;
    int i = 0;
#pragma unroll
    for (int j = 0; j < OCC_MAX_SWAPS - 1; ++j) i += (j < n_swaps - 1 && ts[j] <= r) ? 1 : 0;
    return i;
  };
  // every lane's entry after an accepted event changed the occupant-0 counts
  // of units ua (by -dir_a) and, canonical, ub (by -dir_b)
  auto account = [&](int ua, int dir_a, int ub, int dir_b) {
    if (lane < n_swaps) {
      auto d = [&](int u) { return (u == ua ? -dir_a : 0) + (canonical && u == ub ? -dir_b : 0); };
      mine.na += d(mine.ua) * mine.dir_a;
      mine.nb += d(mine.ub) * mine.dir_b;
    }
  };

  const int base_len = canonical ? 3 : 2;
  const int64_t target = n_passes * (int64_t)n_mut + n_steps;
  int64_t done_steps = 0;
  bool dirty = true;
  while (done_steps < target) {
    if (dirty) {
      publish();
      dirty = false;
    }
    if (Ttot == 0.0) {  // OccEventProposal.hh:136-139 / 290-293: no events possible
      frozen = 1;
      break;
    }
    ensure(pos + W + 3);
    const bool inwin = lane < W;
    const int64_t q = inwin ? pos + lane : pos;
    const uint64_t x0 = word(q), x1 = word(q + 1), x2 = word(q + 2);
    const uint64_t x3 = canonical ? word(q + 3) : 0ULL;
    const double rand = __dmul_rn(canonical_from_word(x0), Ttot);  // random_real(tsum.back())
    const int si = choose(rand);
    Proposal p;
    p.e = tab_s[si];
    const uint64_t na = (uint64_t)p.e.na, nb = (uint64_t)p.e.nb;
    const uint32_t sa = (uint32_t)__umul64hi(x1, na);
    bool special = inwin && (x1 * na) < na;
    p.pa = p.e.base_a + p.e.dir_a * (int32_t)sa;
    p.pba = p.e.base_a + p.e.dir_a * (p.e.na - 1);
    p.a = list[p.pa];
    p.la = list[p.pba];
    uint64_t xm = x2;
    if (canonical) {
      const uint32_t sb = (uint32_t)__umul64hi(x2, nb);
      special = special || (inwin && (x2 * nb) < nb);
      p.pb = p.e.base_b + p.e.dir_b * (int32_t)sb;
      p.pbb = p.e.base_b + p.e.dir_b * (p.e.nb - 1);
      p.b = list[p.pb];
      p.lb = list[p.pbb];
      xm = x3;
    }
    classify(p, si, canonical, lat, rows, cols, magic);
    const uint32_t fl = flg_s[p.k];
    const uint64_t thr = thr_s[p.k];
    const double dq = dq_s[p.k];
    const bool nd = !(fl & 1u);
    const bool acc = !nd || (fl & 2u) || (xm < thr);
    const bool valid = inwin && !special;
    const uint64_t ND = __ballot(valid && nd);
    const uint64_t AC = __ballot(valid && acc);
    const uint64_t SP = __ballot(special);

    // step starts: 0, then +base_len or +base_len+1 (a uniform drawn), up to
    // the first accepted step or before the first special one
    uint64_t exec = 0;
    for (int t = 0; t < W;) {
      const uint64_t bit = 1ull << t;
      if (SP & bit) break;
      exec |= bit;
      if (AC & bit) break;
      t += base_len + (int)((ND >> t) & 1ull);
    }
    int steps = __popcll(exec);
    const int64_t to_pass_end = (int64_t)n_mut - step;
    const int64_t to_target = target - done_steps;
    const int64_t lim = to_pass_end < to_target ? to_pass_end : to_target;
    if (steps > lim) {
      uint64_t rest = exec;
      for (int r = 0; r < lim; ++r) rest &= rest - 1;
      exec ^= rest;
      steps = (int)lim;
    }
    int used = 0;
    if (exec) {
      const int last = 63 - __builtin_clzll(exec);
      used = last + base_len + (int)((ND >> last) & 1ull);
    }
    const uint64_t F = exec & AC;  // at most one step: the last one
    if (F) {
      const int f = __builtin_ctzll(F);
      if (lane == f) apply_proposal(p, canonical, lat, list, cnt_s);
      asm volatile("" ::: "memory");
      e_run = __dadd_rn(e_run, readlane_f64(dq, f));
      S += __builtin_amdgcn_readlane(p.dS, f);
      M += __builtin_amdgcn_readlane(p.dM, f);
      account(__builtin_amdgcn_readlane(p.e.ua, f), __builtin_amdgcn_readlane(p.e.dir_a, f),
              __builtin_amdgcn_readlane(p.e.ub, f), __builtin_amdgcn_readlane(p.e.dir_b, f));
      ++n_acc;
      dirty = true;
    }
    n_rej += steps - (F ? 1 : 0);
    pos += used;
    step += steps;
    done_steps += steps;

    if ((SP & 1ull) && steps == 0 && done_steps < target) {
      // a step whose Lemire draw may reject, with the full rejection loops;
      // every lane computes it, lane 0 writes
      ensure(pos + 1);
      const double r1 = __dmul_rn(canonical_from_word(word(pos)), Ttot);
      ++pos;
      const int i1 = choose(r1);
      Proposal q1;
      q1.e = tab_s[i1];
      const uint32_t s1a = lemire_scalar((uint64_t)q1.e.na);
      q1.pa = q1.e.base_a + q1.e.dir_a * (int32_t)s1a;
      q1.pba = q1.e.base_a + q1.e.dir_a * (q1.e.na - 1);
      q1.a = list[q1.pa];
      q1.la = list[q1.pba];
      if (canonical) {
        const uint32_t s1b = lemire_scalar((uint64_t)q1.e.nb);
        q1.pb = q1.e.base_b + q1.e.dir_b * (int32_t)s1b;
        q1.pbb = q1.e.base_b + q1.e.dir_b * (q1.e.nb - 1);
        q1.b = list[q1.pb];
        q1.lb = list[q1.pbb];
      }
      classify(q1, i1, canonical, lat, rows, cols, magic);
      bool ok;
      if (flg_s[q1.k] & 1u) {
        ok = true;
      } else {
        ensure(pos + 1);
        const uint64_t xr = word(pos);
        ++pos;
        ok = (flg_s[q1.k] & 2u) || (xr < thr_s[q1.k]);
      }
      // one wavefront walks: its LDS operations execute in issue order and the
      // asm keeps the compiler from reordering them
      asm volatile("" ::: "memory");
      if (ok) {
        if (lane == 0) apply_proposal(q1, canonical, lat, list, cnt_s);
        e_run = __dadd_rn(e_run, dq_s[q1.k]);
        S += q1.dS;
        M += q1.dM;
        account(q1.e.ua, q1.e.dir_a, q1.e.ub, q1.e.dir_b);
        ++n_acc;
        dirty = true;
      } else {
        ++n_rej;
      }
      asm volatile("" ::: "memory");
      ++step;
      ++done_steps;
    }

    if (step == n_mut) {
      step = 0;
      ++n_pass;
      if (n_pass % sample_period == 0) {
        if (lane == 0 && n_samples < series_cap) {
          mcg_sample rec = {};
          rec.pass = n_pass;
          rec.S1 = S;
          rec.M = M;
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
  for (int i = lane; i < n; i += 64) g_lat[i] = (int8_t)(lat[i] & 1);
  for (int i = lane; i < n_mut; i += 64) g_list[i] = (int32_t)list[i];
  if (lane < OCC_MAX_ASYM) St->occ_cnt0[lane] = cnt_s[lane];
  if (lane == 0) {
    St->pos = p_fin;
    St->n_pass = n_pass;
    St->n_accept = n_acc;
    St->n_reject = n_rej;
    St->S = S;
    St->M = M;
    St->e_run = e_run;
    St->n_samples = n_samples;
    St->step = step;
    St->frozen = frozen;
  }
}

void launch_occupant(const OccParams *P, const OccTables *T, ExactState *S, int8_t *lat_pool,
                     int32_t *list_pool, mcg_sample *series, int64_t series_cap,
                     const uint8_t *active, int n_chains, int64_t n_passes, int64_t n_steps,
                     int64_t sample_period, int lds_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(occupant_kernel, dim3(n_chains), dim3(128), lds_bytes, stream, P, T, S,
                     lat_pool, list_pool, series, series_cap, active, n_passes, n_steps,
                     sample_period);
}

hipError_t occupant_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(occupant_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
