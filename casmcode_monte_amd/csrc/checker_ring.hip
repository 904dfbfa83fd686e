// checker_ring.hip -- CHECKERBOARD sweeps of lattices with an odd edge.
//
// The split-row kernels (checker.hip) colour a periodic lattice by the parity
// of each coordinate, which needs even edges.  A chain with an odd edge uses a
// product of ring colourings instead: on a ring of n points, point x has
// colour x & 1, except that when n is odd the last point n-1 gets colour 2.
// Points of one colour are then 0 or >= 2 apart around the ring (the wrap gap
// of colours 0 and 1 is 3 when n is odd, 2 when it is even).  A site's class
// is the triple of its coordinates' colours (c0, c1, c2), index c0 + 3 c1 +
// 9 c2; every nearest and next-nearest offset moves some coordinate by exactly
// 1, so two sites of one class never interact and a class updates in parallel
// -- the same argument as the parity colourings.  One pass visits the classes
// in index order (empty ones skipped): every site is attempted once.
//
// Layout: the natural order l = i + L0 (j + L1 k), one byte per site (1 for
// spin +1, 0 for -1).  The sites of class (c0, c1, c2) have coordinates
// x = 2 s + c (colours 0, 1) or x = n - 1 (colour 2); a work unit is up to
// four consecutive class sites s = 4q .. 4q+3 along i of one class line
// (j, k), unit index u = (line index) * ceil(n0 / 4) + q with the class lines
// ordered j fastest.  Its random words are one Philox4x32-10 call with counter
// (u, pass lo, pass hi, class | 2^30) and the chain seed as key, so a site's
// stream depends only on (seed, pass, class, site): the LDS-resident and HBM
// kernels, and a chain alone or in a batch, give identical trajectories.  A
// site flips iff its word r <= t for the chain's threshold table entry
// t = tab[checker_class(b, c1, c2)] (the table of the parity-coloured chains).
#include "internal.hh"

namespace mcg {

namespace {

__device__ __forceinline__ int ring_count(int n, int c) {
  if (c == 2) return n & 1;
  return (n & 1) ? (n - 1) >> 1 : n >> 1;
}

__device__ __forceinline__ int ring_pos(int n, int c, int s) { return c == 2 ? n - 1 : 2 * s + c; }

__device__ __forceinline__ int ring_wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

struct RingClass {
  int c0, c1, c2, col;
  int n0, n1;  // class sites per class line, class lines along j
  int upl;     // work units per class line
  int64_t units;
};

__device__ __forceinline__ RingClass ring_class(const CheckerParams &P, int col) {
  RingClass k;
  k.col = col;
  k.c0 = col % 3;
  k.c1 = (col / 3) % 3;
  k.c2 = col / 9;
  k.n0 = ring_count(P.L0, k.c0);
  k.n1 = ring_count(P.L1, k.c1);
  const int n2 = P.ndim == 3 ? ring_count(P.L2, k.c2) : (k.c2 == 0 ? 1 : 0);
  k.upl = (k.n0 + 3) >> 2;
  k.units = (int64_t)k.n1 * n2 * k.upl;
  return k;
}

__device__ __forceinline__ int ring_n_classes(const CheckerParams &P) { return P.ndim == 3 ? 27 : 9; }

// Update work unit u of class k (lat: the chain's lattice, LDS or HBM).
__device__ __forceinline__ void ring_unit(const CheckerParams &P, const RingClass &k,
                                          uint8_t *__restrict__ lat,
                                          const uint32_t *__restrict__ tab, int64_t u,
                                          int64_t pass, int &dS1, int &dS2, int &dM,
                                          int &nacc) {
  const int L0 = P.L0, L1 = P.L1, L2 = P.L2;
  const int q = (int)(u % k.upl);
  const int64_t li = u / k.upl;
  const int j = ring_pos(L1, k.c1, (int)(li % k.n1));
  const int kz = P.ndim == 3 ? ring_pos(L2, k.c2, (int)(li / k.n1)) : 0;
  const int jp = ring_wrap(j + 1, L1), jm = ring_wrap(j - 1, L1);
  const int64_t plane = (int64_t)L0 * L1;
  const int64_t r0 = kz * plane + (int64_t)j * L0;  // own line
  const int64_t rjp = kz * plane + (int64_t)jp * L0, rjm = kz * plane + (int64_t)jm * L0;
  int64_t rkp = r0, rkm = r0, rjpkp = r0, rjpkm = r0, rjmkp = r0, rjmkm = r0;
  if (P.ndim == 3) {
    const int64_t kp = ring_wrap(kz + 1, L2), km = ring_wrap(kz - 1, L2);
    rkp = kp * plane + (int64_t)j * L0;
    rkm = km * plane + (int64_t)j * L0;
    rjpkp = kp * plane + (int64_t)jp * L0;
    rjpkm = km * plane + (int64_t)jp * L0;
    rjmkp = kp * plane + (int64_t)jm * L0;
    rjmkm = km * plane + (int64_t)jm * L0;
  }
  const u32x4 r = philox4x32_10((uint32_t)u, (uint32_t)pass, (uint32_t)(pass >> 32),
                                (uint32_t)k.col | 0x40000000u, P.key0, P.key1);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int s = 4 * q + v;
    if (s >= k.n0) break;
    const int i = ring_pos(L0, k.c0, s);
    const int ip = ring_wrap(i + 1, L0), im = ring_wrap(i - 1, L0);
    const int b = lat[r0 + i];
    int c1 = lat[r0 + ip] + lat[r0 + im] + lat[rjp + i] + lat[rjm + i];
    int c2 = 0;
    if (P.ndim == 3) c1 += lat[rkp + i] + lat[rkm + i];
    if (P.z2 > 0) {
      c2 = lat[rjp + ip] + lat[rjp + im] + lat[rjm + ip] + lat[rjm + im];
      if (P.ndim == 3)
        c2 += lat[rkp + ip] + lat[rkp + im] + lat[rkm + ip] + lat[rkm + im] + lat[rjpkp + i] +
              lat[rjpkm + i] + lat[rjmkp + i] + lat[rjmkm + i];
    }
    const uint32_t t = tab[checker_class(b, c1, c2)];
    if (r.v[v] <= t) {
      lat[r0 + i] = (uint8_t)(1 - b);
      const int sp = 2 * b - 1;
      dS1 += -2 * sp * (2 * c1 - P.z1);
      dS2 += -2 * sp * (2 * c2 - P.z2);
      dM += -2 * sp;
      nacc += 1;
    }
  }
}

__device__ __forceinline__ int ring_wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One workgroup per chain runs n_passes with the lattice in LDS.
__global__ __launch_bounds__(1024) void checker_ring_lds_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    mcg_sample *__restrict__ series, int64_t series_cap, const uint8_t *__restrict__ active,
    int64_t n_passes, int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int red[4][16];
  __shared__ uint32_t tab[CHECKER_NTAB];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  uint8_t *g_lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  uint8_t *lat = reinterpret_cast<uint8_t *>(smem);
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < CHECKER_NTAB; i += nt) tab[i] = tables[P.tab_off + i];
  for (int64_t i = t; i < P.n; i += nt) lat[i] = g_lat[i];
  __syncthreads();
  int64_t pass = St->n_pass;
  int64_t S1 = St->S1, S2 = St->S2, M = St->M, nacc_tot = St->n_accept;
  int64_t n_samples = St->n_samples;
  const int wid = t >> 6, lane = t & 63, nw = nt >> 6;
  const int ncls = ring_n_classes(P);
  for (int64_t p = 0; p < n_passes; ++p) {
    int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
    for (int col = 0; col < ncls; ++col) {
      const RingClass k = ring_class(P, col);
      if (k.units == 0) continue;  // the same for every thread
      for (int64_t u = t; u < k.units; u += nt)
        ring_unit(P, k, lat, tab, u, pass, dS1, dS2, dM, nacc);
      __syncthreads();
    }
    dS1 = ring_wave_sum(dS1);
    dS2 = ring_wave_sum(dS2);
    dM = ring_wave_sum(dM);
    nacc = ring_wave_sum(nacc);
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
        mcg_sample rec = {};
        rec.pass = pass;
        rec.S1 = S1;
        rec.S2 = S2;
        rec.M = M;
        series[(int64_t)c * series_cap + n_samples] = rec;
      }
      ++n_samples;
    }
  }
  for (int64_t i = t; i < P.n; i += nt) g_lat[i] = lat[i];
  if (t == 0) {
    St->n_pass = pass;
    St->S1 = S1;
    St->S2 = S2;
    St->M = M;
    St->n_accept = nacc_tot;
    St->n_samples = n_samples;
  }
}

// One class of every listed chain on its HBM lattice: block b takes chain
// chains[b / per_chain] and the units (b % per_chain) * 256 + t + m * 256 *
// per_chain; observable changes are added to the chain's state atomically
// (the pass counter and the record follow in checker_record_kernel).
constexpr int RING_BLOCK = 256;
__global__ __launch_bounds__(RING_BLOCK) void checker_ring_class_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ chains, int per_chain, int col) {
  __shared__ int red[4][RING_BLOCK / 64];
  const int c = chains[blockIdx.x / per_chain];
  const int sub = blockIdx.x % per_chain;
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  const RingClass k = ring_class(P, col);
  const int64_t pass = St->n_pass;
  const uint32_t *tab = tables + P.tab_off;
  uint8_t *lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
  const int t = threadIdx.x;
  for (int64_t u = (int64_t)sub * RING_BLOCK + t; u < k.units;
       u += (int64_t)per_chain * RING_BLOCK)
    ring_unit(P, k, lat, tab, u, pass, dS1, dS2, dM, nacc);
  dS1 = ring_wave_sum(dS1);
  dS2 = ring_wave_sum(dS2);
  dM = ring_wave_sum(dM);
  nacc = ring_wave_sum(nacc);
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
    for (int w = 0; w < RING_BLOCK / 64; ++w) {
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

}  // namespace

void launch_checker_ring_lds(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                             const uint32_t *tables, mcg_sample *series, int64_t series_cap,
                             const uint8_t *active, int n_chains, int64_t n_passes,
                             int64_t sample_period, int lds_bytes, int block,
                             hipStream_t stream) {
  hipLaunchKernelGGL(checker_ring_lds_kernel, dim3(n_chains), dim3(block), lds_bytes, stream, P,
                     St, lat_pool, tables, series, series_cap, active, n_passes, sample_period);
}

void launch_checker_ring_pass(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                              const uint32_t *tables, const int32_t *chains, int n_ring,
                              int per_chain, int ndim, hipStream_t stream) {
  const int ncls = ndim == 3 ? 27 : 9;
  for (int col = 0; col < ncls; ++col)
    hipLaunchKernelGGL(checker_ring_class_kernel, dim3(n_ring * per_chain), dim3(RING_BLOCK), 0,
                       stream, P, St, lat_pool, tables, chains, per_chain, col);
}

hipError_t checker_ring_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(checker_ring_lds_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
