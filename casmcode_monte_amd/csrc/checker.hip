// checker.hip -- CHECKERBOARD mode: sublattice-coloured Metropolis sweeps.
//
// Reference algorithm per site: the SGC single-site flip of
// basic_semigrand_canonical.hh:178-192 / model.hh:305-345 with the Metropolis
// rule of metropolis.hh:25-35, generalised to next-nearest couplings and 3D.
// Sites of one colour share no interaction, so all of them are updated at once;
// one pass = one sweep over every colour, i.e. one attempt per site.
//
// Colourings (even edge lengths):  2D NN: (i+j)&1;  2D NN+NNN: (i&1)+2(j&1);
// 3D NN: (i+j+k)&1;  3D NN+NNN: (i&1)+2(j&1)+4(k&1) (the face diagonals of one
// parity sublattice form FCC nearest-neighbour tetrahedra, so 4 colours fail).
//
// Work unit = a "quad": 4 sites of one colour that are consecutive along i in
// one lattice line.  Each quad takes one Philox4x32-10 call keyed by the chain
// seed with counter (quad index, pass lo, pass hi, colour), so the random
// stream of a site depends only on (chain seed, pass, site): the LDS-resident
// kernel, the HBM work-list kernel, a chain run alone and a chain run in a
// batch all produce the same trajectory bit for bit.
//
// Acceptance: class (old spin, NN sum, NNN sum) -> 32-bit threshold t with
// accept iff r <= t (t = 2^32-1 when dE <= 0), tabulated per chain on the host
// with exp(-beta*dE); no exp per attempt.  Observables are exact integers
// (S1 = sum_NN s s, S2 = sum_NNN s s, M = sum s) updated from per-flip deltas
// reduced per workgroup.
#include "internal.hh"

namespace mcg {

struct QuadGeom {
  int32_t half0;  // sites of one colour per line (L0/2)
  int32_t qpl;    // quads per line
  int32_t nl;     // lines per colour
};

__device__ __forceinline__ QuadGeom quad_geom(const CheckerParams &P) {
  QuadGeom g;
  g.half0 = P.L0 >> 1;
  g.qpl = (g.half0 + 3) >> 2;
  if (P.ncolor == 2)
    g.nl = P.L1 * P.L2;
  else if (P.ncolor == 4)
    g.nl = P.L1 >> 1;
  else
    g.nl = (P.L1 >> 1) * (P.L2 >> 1);
  return g;
}

__device__ __forceinline__ int wrapm(int a, int L) { return a < 0 ? a + L : (a >= L ? a - L : a); }

// Update the (up to) 4 sites of quad `qd` of colour `color`.
// Returns per-quad deltas through references.
__device__ __forceinline__ void update_quad(const CheckerParams &P, const QuadGeom &g,
                                            int8_t *__restrict__ lat,
                                            const uint32_t *__restrict__ tab, int color,
                                            int64_t pass, int64_t qd, int &dS1, int &dS2,
                                            int &dM, int &nacc) {
  const int ln = (int)(qd / g.qpl);
  const int qi = (int)(qd - (int64_t)ln * g.qpl);
  int j, k, a;
  if (P.ncolor == 2) {
    j = ln % P.L1;
    k = ln / P.L1;
    a = (color + j + k) & 1;
  } else if (P.ncolor == 4) {
    a = color & 1;
    j = ((color >> 1) & 1) + 2 * ln;
    k = 0;
  } else {
    const int h1 = P.L1 >> 1;
    a = color & 1;
    j = ((color >> 1) & 1) + 2 * (ln % h1);
    k = ((color >> 2) & 1) + 2 * (ln / h1);
  }
  const int L0 = P.L0, L1 = P.L1, L2 = P.L2;
  const int jp = wrapm(j + 1, L1), jm = wrapm(j - 1, L1);
  const int64_t plane = (int64_t)L0 * L1;
  const int64_t kb = (int64_t)k * plane;
  int8_t *row = lat + kb + (int64_t)j * L0;
  const int8_t *rjp = lat + kb + (int64_t)jp * L0;
  const int8_t *rjm = lat + kb + (int64_t)jm * L0;
  const int8_t *rkp = row, *rkm = row;
  const int8_t *rjp_kp = row, *rjp_km = row, *rjm_kp = row, *rjm_km = row;
  if (P.ndim == 3) {
    const int kp = wrapm(k + 1, L2), km = wrapm(k - 1, L2);
    rkp = lat + (int64_t)kp * plane + (int64_t)j * L0;
    rkm = lat + (int64_t)km * plane + (int64_t)j * L0;
    rjp_kp = lat + (int64_t)kp * plane + (int64_t)jp * L0;
    rjp_km = lat + (int64_t)km * plane + (int64_t)jp * L0;
    rjm_kp = lat + (int64_t)kp * plane + (int64_t)jm * L0;
    rjm_km = lat + (int64_t)km * plane + (int64_t)jm * L0;
  }
  const int64_t qglob = (int64_t)ln * g.qpl + qi;
  const u32x4 r = philox4x32_10((uint32_t)qglob, (uint32_t)pass, (uint32_t)(pass >> 32),
                                (uint32_t)color, P.key0, P.key1);
  const int z1 = P.z1, z2 = P.z2;
  const int w2 = z2 + 1;
  const int cls_s = (z1 + 1) * w2;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int ip = 4 * qi + u;
    if (ip >= g.half0) break;
    const int i = a + 2 * ip;
    const int ie = wrapm(i + 1, L0), iw = wrapm(i - 1, L0);
    const int s = row[i];
    int nn = row[ie] + row[iw] + rjp[i] + rjm[i];
    int nnn = 0;
    if (P.ndim == 3) nn += rkp[i] + rkm[i];
    if (z2 > 0) {
      nnn = rjp[ie] + rjp[iw] + rjm[ie] + rjm[iw];
      if (P.ndim == 3)
        nnn += rkp[ie] + rkp[iw] + rkm[ie] + rkm[iw] + rjp_kp[i] + rjp_km[i] + rjm_kp[i] +
               rjm_km[i];
    }
    const int cls = (s > 0 ? cls_s : 0) + ((nn + z1) >> 1) * w2 + ((nnn + z2) >> 1);
    if (r.v[u] <= tab[cls]) {
      row[i] = (int8_t)(-s);
      dS1 += -2 * s * nn;
      dS2 += -2 * s * nnn;
      dM += -2 * s;
      nacc += 1;
    }
  }
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// LDS-resident kernel: one workgroup per chain runs n_passes entirely in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void checker_lds_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    mcg_sample *__restrict__ series, int64_t series_cap, const uint8_t *__restrict__ active,
    int64_t n_passes, int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int red[4][32];
  __shared__ uint32_t tab[2 * 7 * 13];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  int8_t *g_lat = lat_pool + P.lat_off;
  int8_t *lat = reinterpret_cast<int8_t *>(smem);
  const int t = threadIdx.x, nt = blockDim.x;
  const int ntab = 2 * (P.z1 + 1) * (P.z2 + 1);
  for (int i = t; i < ntab; i += nt) tab[i] = tables[P.tab_off + i];
  const int64_t n = P.n;
  if ((n & 15) == 0) {
    const uint4 *src = reinterpret_cast<const uint4 *>(g_lat);
    uint4 *dst = reinterpret_cast<uint4 *>(lat);
    for (int64_t i = t; i < (n >> 4); i += nt) dst[i] = src[i];
  } else {
    for (int64_t i = t; i < n; i += nt) lat[i] = g_lat[i];
  }
  __syncthreads();
  const QuadGeom g = quad_geom(P);
  const int64_t nq = (int64_t)g.nl * g.qpl;
  int64_t pass = St->n_pass;
  int64_t S1 = St->S1, S2 = St->S2, M = St->M, nacc_tot = St->n_accept;
  int64_t n_samples = St->n_samples;
  const int wid = t >> 6, lane = t & 63, nw = nt >> 6;
  for (int64_t p = 0; p < n_passes; ++p) {
    int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
    for (int color = 0; color < P.ncolor; ++color) {
      for (int64_t qd = t; qd < nq; qd += nt)
        update_quad(P, g, lat, tab, color, pass, qd, dS1, dS2, dM, nacc);
      __syncthreads();
    }
    dS1 = wave_sum(dS1);
    dS2 = wave_sum(dS2);
    dM = wave_sum(dM);
    nacc = wave_sum(nacc);
    if (lane == 0) {
      red[0][wid] = dS1;
      red[1][wid] = dS2;
      red[2][wid] = dM;
      red[3][wid] = nacc;
    }
    __syncthreads();
    int64_t a1 = 0, a2 = 0, am = 0, an = 0;
    for (int w = 0; w < nw; ++w) {
      a1 += red[0][w];
      a2 += red[1][w];
      am += red[2][w];
      an += red[3][w];
    }
    __syncthreads();
    S1 += a1;
    S2 += a2;
    M += am;
    nacc_tot += an;
    ++pass;
    if (pass % sample_period == 0) {
      if (t == 0 && n_samples < series_cap) {
        mcg_sample rec;
        rec.pass = pass;
        rec.S1 = S1;
        rec.S2 = S2;
        rec.M = M;
        rec.e_fold = 0.0;
        rec.e_run = 0.0;
        series[(int64_t)c * series_cap + n_samples] = rec;
      }
      ++n_samples;
    }
  }
  if ((n & 15) == 0) {
    const uint4 *src = reinterpret_cast<const uint4 *>(lat);
    uint4 *dst = reinterpret_cast<uint4 *>(g_lat);
    for (int64_t i = t; i < (n >> 4); i += nt) dst[i] = src[i];
  } else {
    for (int64_t i = t; i < n; i += nt) g_lat[i] = lat[i];
  }
  if (t == 0) {
    St->n_pass = pass;
    St->S1 = S1;
    St->S2 = S2;
    St->M = M;
    St->n_accept = nacc_tot;
    St->n_samples = n_samples;
  }
}

void launch_checker_lds(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                        const uint32_t *tables, mcg_sample *series, int64_t series_cap,
                        const uint8_t *active, int n_chains, int64_t n_passes,
                        int64_t sample_period, int lds_bytes, int block, hipStream_t stream) {
  hipLaunchKernelGGL(checker_lds_kernel, dim3(n_chains), dim3(block), lds_bytes, stream, P, St,
                     lat_pool, tables, series, series_cap, active, n_passes, sample_period);
}

void checker_set_max_lds(int bytes) {
  hipFuncSetAttribute(reinterpret_cast<const void *>(checker_lds_kernel),
                      hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// ---------------------------------------------------------------------------
// HBM work-list kernel: one launch per colour over a list of (chain, first
// quad) work items, QUADS_PER_ITEM quads each; the lattice streams from HBM.
// ---------------------------------------------------------------------------
constexpr int HBM_BLOCK = 256;
constexpr int QUADS_PER_THREAD = 4;
constexpr int QUADS_PER_ITEM = HBM_BLOCK * QUADS_PER_THREAD;

__global__ __launch_bounds__(HBM_BLOCK) void checker_color_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ work, int color) {
  __shared__ int red[4][HBM_BLOCK / 64];
  const int c = work[2 * blockIdx.x];
  const int64_t q0 = (int64_t)work[2 * blockIdx.x + 1] * QUADS_PER_ITEM;
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  const QuadGeom g = quad_geom(P);
  const int64_t nq = (int64_t)g.nl * g.qpl;
  const int64_t pass = St->n_pass;
  const uint32_t *tab = tables + P.tab_off;
  int8_t *lat = lat_pool + P.lat_off;
  int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < QUADS_PER_THREAD; ++u) {
    const int64_t qd = q0 + (int64_t)u * HBM_BLOCK + t;
    if (qd < nq) update_quad(P, g, lat, tab, color, pass, qd, dS1, dS2, dM, nacc);
  }
  dS1 = wave_sum(dS1);
  dS2 = wave_sum(dS2);
  dM = wave_sum(dM);
  nacc = wave_sum(nacc);
  const int wid = t >> 6, lane = t & 63;
  if (lane == 0) {
    red[0][wid] = dS1;
    red[1][wid] = dS2;
    red[2][wid] = dM;
    red[3][wid] = nacc;
  }
  __syncthreads();
  if (t == 0) {
    long long a1 = 0, a2 = 0, am = 0, an = 0;
    for (int w = 0; w < HBM_BLOCK / 64; ++w) {
      a1 += red[0][w];
      a2 += red[1][w];
      am += red[2][w];
      an += red[3][w];
    }
    if (a1) atomicAdd(reinterpret_cast<unsigned long long *>(&St->S1), (unsigned long long)a1);
    if (a2) atomicAdd(reinterpret_cast<unsigned long long *>(&St->S2), (unsigned long long)a2);
    if (am) atomicAdd(reinterpret_cast<unsigned long long *>(&St->M), (unsigned long long)am);
    if (an)
      atomicAdd(reinterpret_cast<unsigned long long *>(&St->n_accept), (unsigned long long)an);
  }
}

// After the last colour of a pass: advance the pass counter and append a
// sample record when due.
__global__ void checker_record_kernel(const CheckerParams *__restrict__ Pp,
                                      CheckerState *__restrict__ Sp,
                                      mcg_sample *__restrict__ series, int64_t series_cap,
                                      const uint8_t *__restrict__ active, int n_chains,
                                      int64_t sample_period) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_chains || !active[c]) return;
  CheckerState *St = Sp + c;
  const int64_t pass = St->n_pass + 1;
  St->n_pass = pass;
  if (pass % sample_period == 0) {
    const int64_t ns = St->n_samples;
    if (ns < series_cap) {
      mcg_sample rec;
      rec.pass = pass;
      rec.S1 = St->S1;
      rec.S2 = St->S2;
      rec.M = St->M;
      rec.e_fold = 0.0;
      rec.e_run = 0.0;
      series[(int64_t)c * series_cap + ns] = rec;
    }
    St->n_samples = ns + 1;
  }
}

void launch_checker_color(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                          const uint32_t *tables, const int32_t *work, int n_work_items,
                          int color, int64_t pass, hipStream_t stream) {
  (void)pass;
  hipLaunchKernelGGL(checker_color_kernel, dim3(n_work_items), dim3(HBM_BLOCK), 0, stream, P,
                     St, lat_pool, tables, work, color);
}

void launch_checker_record(const CheckerParams *P, CheckerState *St, mcg_sample *series,
                           int64_t series_cap, const uint8_t *active, int n_chains,
                           int64_t sample_period, hipStream_t stream) {
  hipLaunchKernelGGL(checker_record_kernel, dim3((n_chains + 63) / 64), dim3(64), 0, stream, P,
                     St, series, series_cap, active, n_chains, sample_period);
}

int checker_quads_per_item() { return QUADS_PER_ITEM; }

}  // namespace mcg
