// fcc.hip -- CHECKERBOARD sweeps of the binary FCC cluster expansion.
//
// The model (fcc.hh, include/casm_mi355x/mcgpu.h): occupation sigma in {0,1}
// on FCC sites, pair interactions in shells 1-3 and nearest-neighbour
// triplets.  Flipping site i changes the SGC potential E - mu*N1 by
//   dE = dsigma * (V1 - mu + V2/6 n1 + V3/3 n2 + V4/12 n3 + Vt/8 t),
// dsigma = 1 - 2 sigma_i, with nk the occupied shell-k neighbours and t the
// triangles through i whose two other corners are occupied.  The Metropolis
// probability factorises: exp(-beta dE) = A[sigma][n1][n2] * B[sigma][n3][t],
// both tabulated per chain on the host, so no exp is evaluated per attempt.
// A site flips iff (r + 1/2) 2^-32 < A*B for its 32-bit Philox word r.
//
// Colouring: every doubled coordinate is coloured on its ring with
// fcc_ring_colour (points of one colour are 0 or >= 3 apart); a class is the
// product of the three coordinate colours, so two sites of one class differ
// by >= 3 in some coordinate while every interaction offset is <= 2 in each:
// a class updates in parallel.  One pass visits the 64 classes in order
// col = cx + 4 cy + 16 cz (the 32 with cx+cy+cz odd are empty when every edge
// is even).
//
// Work unit: up to four sites of one class on one lattice line (fixed y, z),
// slots 4q..4q+3 of that line's list of class x values; one Philox4x32-10
// call per unit with counter (line*32 + q, pass lo, pass hi, class) and the
// chain seed as key, so a site's random word depends only on (seed, pass,
// site) and the LDS-resident and global-memory kernels, any block size, and
// a chain alone or in a batch give identical trajectories.
//
// Kernel: one workgroup per chain runs n_passes; the lattice lives in LDS when
// it fits (<= 142 KB) and is read and written in place in HBM otherwise.
// Counts (N1, P1, P2, P3, T) are exact integers updated from per-flip deltas.
#include "fcc.hh"
#include "internal.hh"

namespace mcg {

namespace {

struct FccLists {
  uint8_t y[4][FCC_LIST];      // y values of each colour
  uint8_t z[4][FCC_LIST];      // z values of each colour
  uint8_t xh[4][2][FCC_LIST];  // h = x >> 1 of each colour and parity of x
  int ny[4], nz[4], nx[4][2];
};

__device__ __forceinline__ int fcc_wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wrap_ring(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

struct FccDelta {
  int N1 = 0, P1 = 0, P2 = 0, P3 = 0, T = 0, acc = 0;
};

// Update one work unit (up to 4 sites of class col on line (y, z)).
__device__ __forceinline__ void fcc_unit(uint8_t *__restrict__ lat, const double *__restrict__ tab,
                                         const uint8_t *__restrict__ tri, const FccLists &Ls,
                                         int L0, int L1, int L2, int cx, int cy, int cz, int col,
                                         int maxq, int ny, int u, int64_t pass, uint32_t key0,
                                         uint32_t key1, FccDelta &d) {
  const int q = u % maxq;
  const int li = u / maxq;
  const int yi = li % ny, zi = li / ny;
  const int y = Ls.y[cy][yi], z = Ls.z[cz][zi];
  const int p = (y + z) & 1;
  const int cnt = Ls.nx[cx][p];
  const int s0 = q << 2;
  if (s0 >= cnt) return;
  const int ry = 2 * L1, rz = 2 * L2;
  int ys[5], zs[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    ys[k] = wrap_ring(y + k - 2, ry);
    zs[k] = wrap_ring(z + k - 2, rz) * ry;
  }
  const int line = y + ry * z;
  const u32x4 r = philox4x32_10((uint32_t)(line * 32 + q), (uint32_t)pass,
                                (uint32_t)(pass >> 32), (uint32_t)col, key0, key1);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int s = s0 + v;
    if (s >= cnt) break;
    const int h = Ls.xh[cx][p][s];
    const int hm = h == 0 ? L0 - 1 : h - 1;
    const int hp = h == L0 - 1 ? 0 : h + 1;
    const int hA = p ? h : hm;  // x - 1 on a line of the other parity
    const int hB = p ? hp : h;  // x + 1 on a line of the other parity
#define MCG_FCC_AT(dx, dy, dz)                                                       \
  ((int)lat[(ys[(dy) + 2] + zs[(dz) + 2]) * L0 +                                     \
            ((((dy) + (dz)) & 1) == 0 ? ((dx) == -2 ? hm : ((dx) == 0 ? h : hp))     \
                                      : ((dx) == -1 ? hA : hB))])
    uint32_t mask = 0;
#define MCG_FCC_BIT(k, dx, dy, dz) mask |= (uint32_t)MCG_FCC_AT(dx, dy, dz) << (k);
    MCG_FCC_NN(MCG_FCC_BIT)
    int n2 = 0;
#define MCG_FCC_ADD2(k, dx, dy, dz) n2 += MCG_FCC_AT(dx, dy, dz);
    MCG_FCC_N2(MCG_FCC_ADD2)
    int n3 = 0;
#define MCG_FCC_ADD3(k, dx, dy, dz) n3 += MCG_FCC_AT(dx, dy, dz);
    MCG_FCC_N3(MCG_FCC_ADD3)
#undef MCG_FCC_ADD3
#undef MCG_FCC_ADD2
#undef MCG_FCC_BIT
#undef MCG_FCC_AT
    const int n1 = __popc(mask);
    const int t = tri[mask];
    const int site = line * L0 + h;
    const int sig = lat[site];
    const double pr = tab[sig * FCC_TAB_A + n1 * (FCC_Z2 + 1) + n2] *
                      tab[2 * FCC_TAB_A + sig * FCC_TAB_B + n3 * (FCC_ZT + 1) + t];
    const double uu = ((double)r.v[v] + 0.5) * 0x1p-32;
    if (uu < pr) {
      lat[site] = (uint8_t)(1 - sig);
      const int ds = 1 - 2 * sig;
      d.N1 += ds;
      d.P1 += ds * n1;
      d.P2 += ds * n2;
      d.P3 += ds * n3;
      d.T += ds * t;
      d.acc += 1;
    }
  }
}

template <bool LAT_IN_LDS>
__global__ __launch_bounds__(1024) void fcc_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const double *__restrict__ tables,
    const uint8_t *__restrict__ tri_g, mcg_sample *__restrict__ series, int64_t series_cap,
    const uint8_t *__restrict__ active, int64_t n_passes, int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double tab[FCC_NTAB];
  __shared__ uint8_t tri[4096];
  __shared__ FccLists Ls;
  __shared__ long long red[6][16];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  uint8_t *g_lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < FCC_NTAB; i += nt) tab[i] = tables[P.tab_off + i];
  for (int i = t; i < 4096; i += nt) tri[i] = tri_g[i];
  const int L0 = P.L0, L1 = P.L1, L2 = P.L2;
  if (t < 4) {  // thread a lists the coordinates of colour a on the three rings
    int k = 0;
    for (int v = 0; v < 2 * L1; ++v)
      if (fcc_ring_colour(v, 2 * L1) == t) Ls.y[t][k++] = (uint8_t)v;
    Ls.ny[t] = k;
    k = 0;
    for (int v = 0; v < 2 * L2; ++v)
      if (fcc_ring_colour(v, 2 * L2) == t) Ls.z[t][k++] = (uint8_t)v;
    Ls.nz[t] = k;
    int k0 = 0, k1 = 0;
    for (int v = 0; v < 2 * L0; ++v)
      if (fcc_ring_colour(v, 2 * L0) == t) {
        if (v & 1)
          Ls.xh[t][1][k1++] = (uint8_t)(v >> 1);
        else
          Ls.xh[t][0][k0++] = (uint8_t)(v >> 1);
      }
    Ls.nx[t][0] = k0;
    Ls.nx[t][1] = k1;
  }
  const int64_t n16 = (P.n + 15) >> 4;  // lattices are padded to 256 bytes
  uint8_t *lat;
  if (LAT_IN_LDS) {
    lat = reinterpret_cast<uint8_t *>(smem);
    const uint4 *src = reinterpret_cast<const uint4 *>(g_lat);
    uint4 *dst = reinterpret_cast<uint4 *>(lat);
    for (int64_t i = t; i < n16; i += nt) dst[i] = src[i];
  } else {
    lat = g_lat;
  }
  __syncthreads();
  int64_t pass = St->n_pass;
  int64_t N1 = St->M, P1 = St->S1, P2 = St->S2, P3 = St->S3, T3 = St->T3;
  int64_t nacc_tot = St->n_accept, n_samples = St->n_samples;
  const int wid = t >> 6, lane = t & 63, nw = nt >> 6;
  for (int64_t it = 0; it < n_passes; ++it) {
    FccDelta d;
    for (int col = 0; col < 64; ++col) {
      const int cx = col & 3, cy = (col >> 2) & 3, cz = col >> 4;
      const int ny = Ls.ny[cy], nz = Ls.nz[cz];
      const int mx = Ls.nx[cx][0] > Ls.nx[cx][1] ? Ls.nx[cx][0] : Ls.nx[cx][1];
      const int maxq = (mx + 3) >> 2;
      const int units = ny * nz * maxq;
      if (units == 0) continue;  // the same for every thread
      for (int u = t; u < units; u += nt)
        fcc_unit(lat, tab, tri, Ls, L0, L1, L2, cx, cy, cz, col, maxq, ny, u, pass, P.key0,
                 P.key1, d);
      __syncthreads();
    }
    const int v0 = fcc_wave_sum(d.N1), v1 = fcc_wave_sum(d.P1), v2 = fcc_wave_sum(d.P2);
    const int v3 = fcc_wave_sum(d.P3), v4 = fcc_wave_sum(d.T), v5 = fcc_wave_sum(d.acc);
    if (lane == 0) {
      red[0][wid] = v0;
      red[1][wid] = v1;
      red[2][wid] = v2;
      red[3][wid] = v3;
      red[4][wid] = v4;
      red[5][wid] = v5;
    }
    __syncthreads();
    long long a[6] = {0, 0, 0, 0, 0, 0};
    for (int w = 0; w < nw; ++w)
#pragma unroll
      for (int k = 0; k < 6; ++k) a[k] += red[k][w];
    __syncthreads();
    N1 += a[0];
    P1 += a[1];
    P2 += a[2];
    P3 += a[3];
    T3 += a[4];
    nacc_tot += a[5];
    ++pass;
    if (pass % sample_period == 0) {
      if (t == 0 && n_samples < series_cap) {
        mcg_sample rec = {};
        rec.pass = pass;
        rec.S1 = P1;
        rec.S2 = P2;
        rec.M = N1;
        rec.S3 = P3;
        rec.T3 = T3;
        series[(int64_t)c * series_cap + n_samples] = rec;
      }
      ++n_samples;
    }
  }
  if (LAT_IN_LDS) {
    const uint4 *src = reinterpret_cast<const uint4 *>(lat);
    uint4 *dst = reinterpret_cast<uint4 *>(g_lat);
    for (int64_t i = t; i < n16; i += nt) dst[i] = src[i];
  }
  if (t == 0) {
    St->n_pass = pass;
    St->M = N1;
    St->S1 = P1;
    St->S2 = P2;
    St->S3 = P3;
    St->T3 = T3;
    St->n_accept = nacc_tot;
    St->n_samples = n_samples;
  }
}

}  // namespace

void launch_fcc(const CheckerParams *P, CheckerState *St, int8_t *lat_pool, const double *tables,
                const uint8_t *tri, mcg_sample *series, int64_t series_cap,
                const uint8_t *active, int n_chains, int64_t n_passes, int64_t sample_period,
                int lds_bytes, bool lattice_in_lds, int block, hipStream_t stream) {
  if (lattice_in_lds)
    hipLaunchKernelGGL(fcc_kernel<true>, dim3(n_chains), dim3(block), lds_bytes, stream, P, St,
                       lat_pool, tables, tri, series, series_cap, active, n_passes,
                       sample_period);
  else
    hipLaunchKernelGGL(fcc_kernel<false>, dim3(n_chains), dim3(block), 0, stream, P, St,
                       lat_pool, tables, tri, series, series_cap, active, n_passes,
                       sample_period);
}

hipError_t fcc_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(fcc_kernel<true>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
