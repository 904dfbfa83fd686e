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
// (position = base + dir * slot), and the occupant change of each site
struct SwapTab {
  int32_t na, base_a, dir_a, da;
  int32_t nb, base_b, dir_b, db;
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

  // this lane's swap type (lanes < n_swaps): its candidates' units, occupants
  // and list segments
  int my_ua = 0, my_oa = 0, my_ub = 0, my_ob = 0;
  if (lane < n_swaps) {
    const int ca = P->swap_a[lane], cb = P->swap_b[lane];
    my_ua = P->cand_asym[ca];
    my_oa = P->cand_occ[ca];
    my_ub = P->cand_asym[cb];
    my_ob = P->cand_occ[cb];
  }
  const int32_t my_off_a = P->asym_off[my_ua], my_size_a = P->asym_size[my_ua];
  const int32_t my_off_b = P->asym_off[my_ub], my_size_b = P->asym_size[my_ub];

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

  // the swap choice's cumulative counts tsum[j+1] (exact integers), uniform
  double ts[OCC_MAX_SWAPS];
  double Ttot = 0.0;
  // OccEventProposal.hh:127-139 / 276-288 from the current list sizes; the
  // table of each swap type's lists for the lanes
  auto rebuild = [&]() {
    int64_t w = 0;
    if (lane < n_swaps) {
      const int c0a = cnt_s[my_ua], c0b = cnt_s[my_ub];
      const int na = my_oa ? my_size_a - c0a : c0a;
      const int nb = my_ob ? my_size_b - c0b : c0b;
      SwapTab e;
      e.na = na;
      e.base_a = my_off_a + (my_oa ? my_size_a - 1 : 0);
      e.dir_a = my_oa ? -1 : 1;
      e.da = 1 - 2 * my_oa;  // a unit has two occupants: a transform swaps them
      e.nb = nb;
      e.base_b = my_off_b + (my_ob ? my_size_b - 1 : 0);
      e.dir_b = my_ob ? -1 : 1;
      e.db = 1 - 2 * my_ob;
      tab_s[lane] = e;
      w = canonical ? (int64_t)na * (int64_t)nb : (int64_t)na;
    }
    int64_t run = 0;
#pragma unroll
    for (int j = 0; j < OCC_MAX_SWAPS; ++j) {
      const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)w, j);
      const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)w >> 32), j);
      run += (int64_t)(((uint64_t)hi << 32) | lo);
      ts[j] = (double)run;
    }
    Ttot = (double)run;
    asm volatile("" ::: "memory");
  };
  // first j with rand < tsum[j+1]: the number of tsum[j+1] <= rand (they do
  // not decrease, and rand < tsum[n_swaps])
  auto choose = [&](double rand) -> int {
    int i = 0;
#pragma unroll
    for (int j = 0; j < OCC_MAX_SWAPS - 1; ++j) i += (j < n_swaps - 1 && ts[j] <= rand) ? 1 : 0;
    return i;
  };

  const int base_len = canonical ? 3 : 2;
  const int64_t target = n_passes * (int64_t)n_mut + n_steps;
  int64_t done_steps = 0;
  bool dirty = true;
  while (done_steps < target) {
    if (dirty) {
      rebuild();
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
    const SwapTab e = tab_s[si];
    const uint64_t na = (uint64_t)e.na, nb = (uint64_t)e.nb;
    const uint32_t sa = (uint32_t)__umul64hi(x1, na);
    bool special = inwin && (x1 * na) < na;
    const int32_t pa = e.base_a + e.dir_a * (int32_t)sa;
    const uint32_t a = list[pa];
    const int oa = lat[a] & 1;
    const int sa_new = 2 * (oa + e.da) - 1;
    const Nbrs ma = nbrs_of(a, rows, cols, magic);
    int Sa, Sa_ex;
    nbr_sums(lat, a, ma, 0xFFFFFFFFu, 0, Sa, Sa_ex);
    int k = si * 25 + ((Sa + 4) >> 1) * 5;
    int dS = (sa_new - (2 * oa - 1)) * Sa_ex, dM = sa_new - (2 * oa - 1);
    uint32_t b = 0;
    int32_t pb = 0;
    uint32_t sb = 0;
    uint64_t xm = x2;
    if (canonical) {
      sb = (uint32_t)__umul64hi(x2, nb);
      special = special || (inwin && (x2 * nb) < nb);
      pb = e.base_b + e.dir_b * (int32_t)sb;
      b = list[pb];
      const int ob = lat[b] & 1;
      const int sb_new = 2 * (ob + e.db) - 1;
      const Nbrs mb = nbrs_of(b, rows, cols, magic);
      int Sb, Sb_ex;
      nbr_sums(lat, b, mb, a, sa_new, Sb, Sb_ex);
      k += (Sb + 4) >> 1;
      dS += (sb_new - (2 * ob - 1)) * Sb_ex;
      dM += sb_new - (2 * ob - 1);
      xm = x3;
    }
    const uint32_t fl = flg_s[k];
    const bool nd = !(fl & 1u);
    const bool acc = !nd || (fl & 2u) || (xm < thr_s[k]);
    const bool valid = inwin && !special;
    const uint64_t ND = __ballot(valid && nd);
    const uint64_t AC = __ballot(valid && acc);
    const uint64_t SP = __ballot(special);

    // step starts: 0, then +base_len or +base_len+1 (a uniform drawn); the
    // block runs to the first accepted step or the first special one
    uint64_t A = 0;
    for (int t = 0; t < W; t += base_len + (int)((ND >> t) & 1ull)) A |= 1ull << t;
    const uint64_t stops_sp = A & SP;
    const uint64_t accs = A & AC & ~SP;
    uint64_t exec = A;
    if (stops_sp) exec &= (1ull << __builtin_ctzll(stops_sp)) - 1ull;
    if (accs) {
      const int fa = __builtin_ctzll(accs);
      exec &= fa == 63 ? ~0ull : ((2ull << fa) - 1ull);
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
      const int fl_lane = __builtin_ctzll(F);
      if (lane == fl_lane) {
        // OccLocation.cc:254-283 for a, then for b (the second transform sees
        // the first's lists)
        const int ua = P->cand_asym[P->swap_a[si]];
        asm volatile("" ::: "memory");
        const int32_t back_a = e.base_a + e.dir_a * (int32_t)(e.na - 1);
        list[pa] = list[back_a];
        asm volatile("" ::: "memory");
        list[back_a] = (uint16_t)a;
        lat[a] = (uint8_t)(oa + e.da);
        cnt_s[ua] += oa == 0 ? -1 : 1;
        if (canonical) {
          const int ub = P->cand_asym[P->swap_b[si]];
          const int32_t nb2 = e.nb + (ub == ua ? 1 : 0);
          const int32_t back_b = e.base_b + e.dir_b * (nb2 - 1);
          const int ob = lat[b] & 1;
          asm volatile("" ::: "memory");
          list[pb] = list[back_b];
          asm volatile("" ::: "memory");
          list[back_b] = (uint16_t)b;
          lat[b] = (uint8_t)(ob + e.db);
          cnt_s[ub] += ob == 0 ? -1 : 1;
        }
      }
      asm volatile("" ::: "memory");
      e_run = __dadd_rn(e_run, readlane_f64(dq_s[k], fl_lane));
      S += __builtin_amdgcn_readlane(dS, fl_lane);
      M += __builtin_amdgcn_readlane(dM, fl_lane);
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
      const SwapTab e1 = tab_s[i1];
      const uint32_t s1a = lemire_scalar((uint64_t)e1.na);
      const int32_t p1a = e1.base_a + e1.dir_a * (int32_t)s1a;
      const uint32_t a1 = list[p1a];
      const int o1a = lat[a1] & 1;
      const int s1a_new = 2 * (o1a + e1.da) - 1;
      int Sa1, Sa1_ex;
      nbr_sums(lat, a1, nbrs_of(a1, rows, cols, magic), 0xFFFFFFFFu, 0, Sa1, Sa1_ex);
      int k1 = i1 * 25 + ((Sa1 + 4) >> 1) * 5;
      int dS1 = (s1a_new - (2 * o1a - 1)) * Sa1_ex, dM1 = s1a_new - (2 * o1a - 1);
      uint32_t b1 = 0;
      int32_t p1b = 0;
      int o1b = 0;
      if (canonical) {
        const uint32_t s1b = lemire_scalar((uint64_t)e1.nb);
        p1b = e1.base_b + e1.dir_b * (int32_t)s1b;
        b1 = list[p1b];
        o1b = lat[b1] & 1;
        const int s1b_new = 2 * (o1b + e1.db) - 1;
        int Sb1, Sb1_ex;
        nbr_sums(lat, b1, nbrs_of(b1, rows, cols, magic), a1, s1a_new, Sb1, Sb1_ex);
        k1 += (Sb1 + 4) >> 1;
        dS1 += (s1b_new - (2 * o1b - 1)) * Sb1_ex;
        dM1 += s1b_new - (2 * o1b - 1);
      }
      bool ok;
      if (flg_s[k1] & 1u) {
        ok = true;
      } else {
        ensure(pos + 1);
        const uint64_t xr = word(pos);
        ++pos;
        ok = (flg_s[k1] & 2u) || (xr < thr_s[k1]);
      }
      // one wavefront walks: its LDS operations execute in issue order and the
      // asm keeps the compiler from reordering them
      asm volatile("" ::: "memory");
      if (ok) {
        const int ua = P->cand_asym[P->swap_a[i1]];
        if (lane == 0) {
          const int32_t back_a = e1.base_a + e1.dir_a * (int32_t)(e1.na - 1);
          list[p1a] = list[back_a];
          asm volatile("" ::: "memory");
          list[back_a] = (uint16_t)a1;
          lat[a1] = (uint8_t)(o1a + e1.da);
          cnt_s[ua] += o1a == 0 ? -1 : 1;
          if (canonical) {
            const int ub = P->cand_asym[P->swap_b[i1]];
            const int32_t nb2 = e1.nb + (ub == ua ? 1 : 0);
            const int32_t back_b = e1.base_b + e1.dir_b * (nb2 - 1);
            asm volatile("" ::: "memory");
            list[p1b] = list[back_b];
            asm volatile("" ::: "memory");
            list[back_b] = (uint16_t)b1;
            lat[b1] = (uint8_t)(o1b + e1.db);
            cnt_s[ub] += o1b == 0 ? -1 : 1;
          }
        }
        e_run = __dadd_rn(e_run, dq_s[k1]);
        S += dS1;
        M += dM1;
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
