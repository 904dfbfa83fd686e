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
// fcc_ring_colour (points of one colour are 0 or >= 3 apart; six colours, the
// last two only on the seams of odd rings); a class is the product of the
// three coordinate colours, col = cx + 6 cy + 36 cz, so two sites of one class
// differ by >= 3 in some coordinate while every interaction offset is <= 2 in
// each: a class updates in parallel.  One pass sweeps the chain's class
// schedule (FccSched, fcc.hh): steps of classes that do not interact, each
// step updated at once, one barrier per step.
//
// Work unit: up to four sites of one class on one lattice line (fixed y, z),
// slots 4q..4q+3 of that line's list of class x values; one Philox4x32-10
// call per unit with counter (line*32 + q, pass lo, pass hi, class) and the
// chain seed as key, so a site's random word depends only on (seed, pass,
// site) and the LDS-resident and global-memory kernels, any block size, and
// a chain alone or in a batch give identical trajectories.
//
// Kernel: one workgroup per chain runs n_passes; the lattice lives in LDS when
// it fits (<= 142 KB, lines padded with ghost copies of their end sites) and
// is read and written in place in HBM otherwise.  A thread gathers the
// neighbourhoods of its unit's four sites before flipping any of them.
// Counts (N1, P1, P2, P3, T) are exact integers updated from per-flip deltas.
#include "fcc.hh"
#include "internal.hh"

namespace mcg {

namespace {

struct FccLists {
  uint8_t y[FCC_NCOL][2][FCC_LIST];   // y values of each colour, by parity
  uint8_t z[FCC_NCOL][2][FCC_LIST];   // z values of each colour, by parity
  uint8_t xh[FCC_NCOL][2][FCC_LIST];  // h = x >> 1 of each colour, by parity of x
  int ny[FCC_NCOL][2], nz[FCC_NCOL][2], nx[FCC_NCOL][2];
};

__device__ __forceinline__ int fcc_wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wrap_ring(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Thread a (< FCC_NCOL) lists the coordinates of colour a on the three rings.
__device__ __forceinline__ void fcc_fill_lists(FccLists &Ls, int a, int L0, int L1, int L2) {
  int k0 = 0, k1 = 0;
  for (int v = 0; v < 2 * L1; ++v)
    if (fcc_ring_colour(v, 2 * L1) == a) {
      if (v & 1) Ls.y[a][1][k1++] = (uint8_t)v;
      else Ls.y[a][0][k0++] = (uint8_t)v;
    }
  Ls.ny[a][0] = k0;
  Ls.ny[a][1] = k1;
  k0 = k1 = 0;
  for (int v = 0; v < 2 * L2; ++v)
    if (fcc_ring_colour(v, 2 * L2) == a) {
      if (v & 1) Ls.z[a][1][k1++] = (uint8_t)v;
      else Ls.z[a][0][k0++] = (uint8_t)v;
    }
  Ls.nz[a][0] = k0;
  Ls.nz[a][1] = k1;
  k0 = k1 = 0;
  for (int v = 0; v < 2 * L0; ++v)
    if (fcc_ring_colour(v, 2 * L0) == a) {
      if (v & 1) Ls.xh[a][1][k1++] = (uint8_t)(v >> 1);
      else Ls.xh[a][0][k0++] = (uint8_t)(v >> 1);
    }
  Ls.nx[a][0] = k0;
  Ls.nx[a][1] = k1;
}

struct FccDelta {
  int N1 = 0, P1 = 0, P2 = 0, P3 = 0, T = 0, acc = 0;
};

// Work units of one class: four blocks (py, pz) of lines whose y and z have
// parities py, pz, each line holding ceil(nx[cx][py^pz] / 4) units.
struct FccClass {
  int cx, cy, cz, col;
  int start[5];  // unit prefix sums of the four blocks
  int nq[4];     // units per line in each block
};

__device__ __forceinline__ FccClass fcc_class(const FccLists &Ls, int col) {
  FccClass k;
  k.col = col;
  k.cx = col % FCC_NCOL;
  k.cy = (col / FCC_NCOL) % FCC_NCOL;
  k.cz = col / (FCC_NCOL * FCC_NCOL);
  k.start[0] = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int py = b >> 1, pz = b & 1;
    k.nq[b] = (Ls.nx[k.cx][py ^ pz] + 3) >> 2;
    k.start[b + 1] = k.start[b] + Ls.ny[k.cy][py] * Ls.nz[k.cz][pz] * k.nq[b];
  }
  return k;
}

// Lattice access.  LDS: every line is stored as [site L0-1, sites 0..L0-1,
// site 0] (ghost copies of the two ends) with a stride that spreads the lines
// of one colour over the banks, so a neighbour at h + dh (|dh| <= 1) is a
// constant offset.  Global memory: plain lines of L0 sites, h wrapped.
template <bool LDS>
struct FccLat {
  uint8_t *p;
  int stride, L0;
  __device__ __forceinline__ int base(int line) const { return line * stride + (LDS ? 1 : 0); }
  __device__ __forceinline__ int at(int b, int h, int dh) const {
    if (LDS) return p[b + h + dh];
    int x = h + dh;
    x = x < 0 ? x + L0 : (x >= L0 ? x - L0 : x);
    return p[b + x];
  }
  __device__ __forceinline__ void flip(int b, int h, uint8_t v) const {
    p[b + h] = v;
    if (LDS) {
      if (h == 0) p[b + L0] = v;       // right ghost of the line
      if (h == L0 - 1) p[b - 1] = v;   // left ghost of the line
    }
  }
};

// Update one work unit: up to 4 sites of class k on one line.  The sites do
// not interact, so the neighbourhoods of a group of FCC_GATHER sites are all
// read before any of them flips (independent loads in flight).
constexpr int FCC_GATHER = 1;
template <bool LDS>
__device__ __forceinline__ void fcc_unit(const FccLat<LDS> &lat, const double *__restrict__ tab,
                                         const uint8_t *__restrict__ tri, const FccLists &Ls,
                                         int L1, int L2, const FccClass &k, int u, int64_t pass,
                                         uint32_t key0, uint32_t key1, FccDelta &d) {
  const int b = (u >= k.start[1]) + (u >= k.start[2]) + (u >= k.start[3]);
  const int py = b >> 1, pz = b & 1, p = py ^ pz;
  const int r0 = u - (b == 0 ? 0 : b == 1 ? k.start[1] : b == 2 ? k.start[2] : k.start[3]);
  const int nq = b == 0 ? k.nq[0] : b == 1 ? k.nq[1] : b == 2 ? k.nq[2] : k.nq[3];
  const int q = r0 % nq;
  const int li = r0 / nq;
  const int nyb = Ls.ny[k.cy][py];
  const int y = Ls.y[k.cy][py][li % nyb], z = Ls.z[k.cz][pz][li / nyb];
  const int cnt = Ls.nx[k.cx][p];
  const int s0 = q << 2;
  const int ry = 2 * L1, rz = 2 * L2;
  int ys[5], zs[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    ys[j] = wrap_ring(y + j - 2, ry);
    zs[j] = wrap_ring(z + j - 2, rz) * ry;
  }
  const int line = y + ry * z;
  const u32x4 r = philox4x32_10((uint32_t)(line * 32 + q), (uint32_t)pass,
                                (uint32_t)(pass >> 32), (uint32_t)k.col, key0, key1);
  const int bl = lat.base(line);
  // sites in groups of FCC_GATHER: gather their neighbourhoods, then decide
#pragma unroll
  for (int g0 = 0; g0 < 4; g0 += FCC_GATHER) {
    uint32_t mask[FCC_GATHER];
    int n2[FCC_GATHER], n3[FCC_GATHER], sig[FCC_GATHER], hh[FCC_GATHER];
#pragma unroll
    for (int j = 0; j < FCC_GATHER; ++j) {
      const int v = g0 + j;
      const int s = s0 + v < cnt ? s0 + v : cnt - 1;  // tail slots repeat the last site
      const int h = Ls.xh[k.cx][p][s];
      const int hq = h + p;  // anchor of the odd-(dy+dz) offsets: x -+ 1 lands at hq-1, hq
      hh[j] = h;
#define MCG_FCC_AT(dx, dy, dz)                                                           \
  lat.at(lat.base(ys[(dy) + 2] + zs[(dz) + 2]), ((((dy) + (dz)) & 1) == 0) ? h : hq,     \
         ((((dy) + (dz)) & 1) == 0) ? (dx) / 2 : ((dx) == -1 ? -1 : 0))
      uint32_t m = 0;
#define MCG_FCC_BIT(kk, dx, dy, dz) m |= (uint32_t)MCG_FCC_AT(dx, dy, dz) << (kk);
      MCG_FCC_NN(MCG_FCC_BIT)
      int a2 = 0;
#define MCG_FCC_ADD2(kk, dx, dy, dz) a2 += MCG_FCC_AT(dx, dy, dz);
      MCG_FCC_N2(MCG_FCC_ADD2)
      int a3 = 0;
#define MCG_FCC_ADD3(kk, dx, dy, dz) a3 += MCG_FCC_AT(dx, dy, dz);
      MCG_FCC_N3(MCG_FCC_ADD3)
#undef MCG_FCC_ADD3
#undef MCG_FCC_ADD2
#undef MCG_FCC_BIT
#undef MCG_FCC_AT
      mask[j] = m;
      n2[j] = a2;
      n3[j] = a3;
      sig[j] = lat.at(bl, h, 0);
    }
#pragma unroll
    for (int j = 0; j < FCC_GATHER; ++j) {
      const int v = g0 + j;
      const int n1 = __popc(mask[j]);
      const int t = tri[mask[j]];
      const int sg = sig[j];
      const double pr = tab[sg * FCC_TAB_A + n1 * (FCC_Z2 + 1) + n2[j]] *
                        tab[2 * FCC_TAB_A + sg * FCC_TAB_B + n3[j] * (FCC_ZT + 1) + t];
      const double uu = ((double)r.v[v] + 0.5) * 0x1p-32;
      if (s0 + v < cnt && uu < pr) {
        lat.flip(bl, hh[j], (uint8_t)(1 - sg));
        const int ds = 1 - 2 * sg;
        d.N1 += ds;
        d.P1 += ds * n1;
        d.P2 += ds * n2[j];
        d.P3 += ds * n3[j];
        d.T += ds * t;
        d.acc += 1;
      }
    }
  }
}

// The class schedule of chain c into LDS, with the prefix sums over its
// classes (in sweep order) of their work units (4-site units: UNITS, or
// class lines); needs Ls filled and a barrier after.
template <bool UNITS>
__device__ __forceinline__ void fcc_sched_setup(const FccSched *__restrict__ sched, int c,
                                                FccSched &S, int *item_start,
                                                const FccLists &Ls, int t) {
  if (t == 0) {
    S = sched[c];
    int acc = 0;
    for (int k = 0; k < S.n_classes; ++k) {
      item_start[k] = acc;
      const FccClass kc = fcc_class(Ls, S.cls[k]);
      if (UNITS) {
        acc += kc.start[4];
      } else {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int py = b >> 1, pz = b & 1;
          acc += Ls.nx[kc.cx][py ^ pz] > 0 ? Ls.ny[kc.cy][py] * Ls.nz[kc.cz][pz] : 0;
        }
      }
    }
    item_start[S.n_classes] = acc;
  }
}

// the position (in sweep order) of the class holding item u of step k0..k1
__device__ __forceinline__ int fcc_item_class(const int *item_start, int k0, int k1, int u) {
  int k = k0;
  while (k + 1 < k1 && u >= item_start[k + 1]) ++k;
  return k;
}

template <bool LAT_IN_LDS>
__global__ __launch_bounds__(1024) void fcc_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const double *__restrict__ tables,
    const uint8_t *__restrict__ tri_g, const FccSched *__restrict__ sched,
    mcg_sample *__restrict__ series, int64_t series_cap,
    const uint8_t *__restrict__ active, int64_t n_passes, int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double tab[FCC_NTAB];
  __shared__ uint8_t tri[4096];
  __shared__ FccLists Ls;
  __shared__ FccSched S;
  __shared__ int item_start[FCC_SCHED_MAX + 1];
  __shared__ long long red[6][16];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  uint8_t *g_lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < FCC_NTAB; i += nt) tab[i] = tables[P.tab_off + i];
  for (int i = t; i < 4096; i += nt) tri[i] = tri_g[i];
  const int L0 = P.L0, L1 = P.L1, L2 = P.L2;
  if (t < FCC_NCOL) fcc_fill_lists(Ls, t, L0, L1, L2);
  __syncthreads();
  fcc_sched_setup<true>(sched, c, S, item_start, Ls, t);
  // In LDS every line takes L0 + 2 bytes (+1 when that is a multiple of 4)
  // with ghost copies of its end sites: the lines a wavefront reads at once
  // are 4 apart (one colour), and this stride puts them in distinct banks.
  const int n_lines = 4 * L1 * L2;
  FccLat<LAT_IN_LDS> lat;
  lat.L0 = L0;
  if (LAT_IN_LDS) {
    const int stride = fcc_lds_stride(L0);
    lat.p = reinterpret_cast<uint8_t *>(smem);
    lat.stride = stride;
    for (int ln = t; ln < n_lines; ln += nt) {
      uint8_t *dst = lat.p + ln * stride;
      const uint8_t *src = g_lat + ln * L0;
      for (int h = 0; h < L0; ++h) dst[1 + h] = src[h];
      dst[0] = src[L0 - 1];
      dst[L0 + 1] = src[0];
    }
  } else {
    lat.p = g_lat;
    lat.stride = L0;
  }
  __syncthreads();
  int64_t pass = St->n_pass;
  int64_t N1 = St->M, P1 = St->S1, P2 = St->S2, P3 = St->S3, T3 = St->T3;
  int64_t nacc_tot = St->n_accept, n_samples = St->n_samples;
  const int wid = t >> 6, lane = t & 63, nw = nt >> 6;
  const int n_steps = S.n_steps;
  for (int64_t it = 0; it < n_passes; ++it) {
    FccDelta d;
    for (int s = 0; s < n_steps; ++s) {
      const int k0 = S.start[s], k1 = S.start[s + 1];
      for (int u = item_start[k0] + t; u < item_start[k1]; u += nt) {
        const int k = fcc_item_class(item_start, k0, k1, u);
        const FccClass kc = fcc_class(Ls, S.cls[k]);
        fcc_unit<LAT_IN_LDS>(lat, tab, tri, Ls, L1, L2, kc, u - item_start[k], pass, P.key0,
                             P.key1, d);
      }
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
  if (LAT_IN_LDS)
    for (int ln = t; ln < n_lines; ln += nt)
      for (int h = 0; h < L0; ++h) g_lat[ln * L0 + h] = lat.p[ln * lat.stride + 1 + h];
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

// ---------------------------------------------------------------------------
// Bit-line kernel: LDS-resident chains with L0 <= 64.
//
// Each lattice line (fixed y, z) is one 64-bit word D = core | core << L0,
// bit h of core the occupation of site h, so the three sites h-1, h, h+1 of
// a line -- with the wrap -- are bits [st, st + 3) of D, st = h - 1 (mod L0),
// one alignbit.  A site's 42 neighbours lie on 21 lines: the even lines
// (dy + dz even: the site's own line, (+-1, +-1), (+-2, 0), (0, +-2)) hold
// them at h - 1, h, h + 1 (dx = -2, 0, +2), the odd lines ((+-1, 0),
// (0, +-1), (+-2, +-1), (+-1, +-2)) at h + p - 1, h + p (dx = -1, +1), p the
// parity of y + z.  A unit's four sites share their 21 lines: 21 word reads
// per unit instead of one byte read per neighbour.  Lines of other classes
// never hold sites of the class being updated (points of one colour are 0 or
// >= 3 apart in every coordinate), so the only concurrent writes a unit sees
// are other units' flips in its own line, whose bits it does not read; a
// unit's flips go into its line with one 64-bit atomic xor.  Units, random
// words and decisions are those of fcc_kernel: trajectories are identical.
// Words are stored at ln + ln / 4, so the lines four apart that a wavefront
// reads together fall into distinct banks.
//
// Lines of 33 to 64 sites (WIDE): the word holds the line once (bits 0..L0-1)
// and the doubled line is formed in registers when a line is read, as the
// 96 bits lo = core | core << L0 (bits 0..63) and hi = core >> (64 - L0)
// (bits 64..95); a window starting at st < L0 <= 64 then lies in lo for
// st < 32 and in the top half of lo and hi otherwise.
__device__ __forceinline__ int fcc_word_index(int ln) { return ln + (ln >> 2); }

__device__ __forceinline__ uint32_t fcc_window(uint64_t D, int st) {
  return __builtin_amdgcn_alignbit((uint32_t)(D >> 32), (uint32_t)D, (uint32_t)st);
}

// one line as the windows read it: the doubled line (narrow: the stored word
// itself; WIDE: lo and hi formed from the stored line)
template <bool WIDE>
struct FccLine {
  uint64_t lo;
  uint32_t hi;
};

template <bool WIDE>
__device__ __forceinline__ FccLine<WIDE> fcc_line(uint64_t stored, int L0) {
  FccLine<WIDE> l;
  if (WIDE) {
    l.lo = L0 < 64 ? stored | (stored << L0) : stored;
    l.hi = (uint32_t)(stored >> (64 - L0));
  } else {
    l.lo = stored;
    l.hi = 0u;
  }
  return l;
}

template <bool WIDE>
__device__ __forceinline__ uint32_t fcc_win(const FccLine<WIDE> &l, int st) {
  if (!WIDE) return fcc_window(l.lo, st);
  const uint32_t a = (uint32_t)l.lo, b = (uint32_t)(l.lo >> 32);
  return st < 32 ? __builtin_amdgcn_alignbit(b, a, (uint32_t)st)
                 : __builtin_amdgcn_alignbit(l.hi, b, (uint32_t)(st - 32));
}

template <bool WIDE>
__device__ __forceinline__ void fcc_bits_unit(uint64_t *__restrict__ w,
                                              const double *__restrict__ tab,
                                              const uint8_t *__restrict__ tri, const FccLists &Ls,
                                              int L0, int L1, int L2, const FccClass &k, int u,
                                              int64_t pass, uint32_t key0, uint32_t key1,
                                              FccDelta &d) {
  const int b = (u >= k.start[1]) + (u >= k.start[2]) + (u >= k.start[3]);
  const int py = b >> 1, pz = b & 1, p = py ^ pz;
  const int r0 = u - (b == 0 ? 0 : b == 1 ? k.start[1] : b == 2 ? k.start[2] : k.start[3]);
  const int nq = b == 0 ? k.nq[0] : b == 1 ? k.nq[1] : b == 2 ? k.nq[2] : k.nq[3];
  const int q = r0 % nq;
  const int li = r0 / nq;
  const int nyb = Ls.ny[k.cy][py];
  const int y = Ls.y[k.cy][py][li % nyb], z = Ls.z[k.cz][pz][li / nyb];
  const int cnt = Ls.nx[k.cx][p];
  const int s0 = q << 2;
  const int ry = 2 * L1, rz = 2 * L2;
  int ys[5], zs[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    ys[j] = wrap_ring(y + j - 2, ry);
    zs[j] = wrap_ring(z + j - 2, rz) * ry;
  }
  typedef FccLine<WIDE> Line;
  auto line = [&](int dy, int dz) -> Line {
    return fcc_line<WIDE>(w[fcc_word_index(ys[dy + 2] + zs[dz + 2])], L0);
  };
  const int own = fcc_word_index(y + ry * z);
  // the unit's four sites (tail slots repeat the last site): h, the even
  // window start se (h-1, h, h+1) and the odd one so (h+p-1, h+p)
  int hh[4], se[4], so[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int s = s0 + v < cnt ? s0 + v : cnt - 1;
    hh[v] = Ls.xh[k.cx][p][s];
    se[v] = hh[v] == 0 ? L0 - 1 : hh[v] - 1;
    so[v] = hh[v] + p - 1 + (hh[v] + p - 1 < 0 ? L0 : 0);
  }
  // the counts of the four sites, line by line: own spin sg, second shell n2,
  // third shell n3, and the nearest-neighbour mask in the bit order of the
  // host's fcc_tri_table_lines (bits 0-7 the odd lines (-1,0), (1,0), (0,-1),
  // (0,1) at x-1, x+1; bits 8-11 the even lines (-1,-1), (-1,1), (1,-1), (1,1))
  int sg[4], n2[4], n3[4];
  uint32_t mask[4];
  {
    const Line L = fcc_line<WIDE>(w[own], L0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t ws = fcc_win<WIDE>(L, se[v]);
      sg[v] = (ws >> 1) & 1;
      n2[v] = __popc(ws & 5u);  // x -+ 2 on the own line
      n3[v] = 0;
      mask[v] = 0u;
    }
  }
  auto even2 = [&](int dy, int dz) {  // second shell: (0, +-2, 0), (0, 0, +-2)
    const Line L = line(dy, dz);
#pragma unroll
    for (int v = 0; v < 4; ++v) n2[v] += (int)((fcc_win<WIDE>(L, se[v]) >> 1) & 1u);
  };
  even2(-2, 0);
  even2(2, 0);
  even2(0, -2);
  even2(0, 2);
  auto even1 = [&](int dy, int dz, int bit) {  // nearest (dx = 0) and third (dx = +-2)
    const Line L = line(dy, dz);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t wv = fcc_win<WIDE>(L, se[v]);
      n3[v] += __popc(wv & 5u);
      mask[v] |= ((wv >> 1) & 1u) << bit;
    }
  };
  even1(-1, -1, 8);
  even1(-1, 1, 9);
  even1(1, -1, 10);
  even1(1, 1, 11);
  auto odd3 = [&](int dy, int dz) {  // third shell (dx = +-1)
    const Line L = line(dy, dz);
#pragma unroll
    for (int v = 0; v < 4; ++v) n3[v] += __popc(fcc_win<WIDE>(L, so[v]) & 3u);
  };
  odd3(-2, -1);
  odd3(-2, 1);
  odd3(2, -1);
  odd3(2, 1);
  odd3(-1, -2);
  odd3(1, -2);
  odd3(-1, 2);
  odd3(1, 2);
  auto odd1 = [&](int dy, int dz, int shift) {  // nearest (dx = +-1)
    const Line L = line(dy, dz);
#pragma unroll
    for (int v = 0; v < 4; ++v) mask[v] |= (fcc_win<WIDE>(L, so[v]) & 3u) << shift;
  };
  odd1(-1, 0, 0);
  odd1(1, 0, 2);
  odd1(0, -1, 4);
  odd1(0, 1, 6);
  const u32x4 r = philox4x32_10((uint32_t)((y + ry * z) * 32 + q), (uint32_t)pass,
                                (uint32_t)(pass >> 32), (uint32_t)k.col, key0, key1);
  uint64_t flips = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int n1 = __popc(mask[v]);
    const int t = tri[mask[v]];
    const double pr = tab[sg[v] * FCC_TAB_A + n1 * (FCC_Z2 + 1) + n2[v]] *
                      tab[2 * FCC_TAB_A + sg[v] * FCC_TAB_B + n3[v] * (FCC_ZT + 1) + t];
    const double uu = ((double)r.v[v] + 0.5) * 0x1p-32;
    if (s0 + v < cnt && uu < pr) {
      const int h = hh[v];
      flips |= WIDE ? (1ull << h) : ((1ull << h) | (1ull << (h + L0)));
      const int ds = 1 - 2 * sg[v];
      d.N1 += ds;
      d.P1 += ds * n1;
      d.P2 += ds * n2[v];
      d.P3 += ds * n3[v];
      d.T += ds * t;
      d.acc += 1;
    }
  }
  if (flips) atomicXor(reinterpret_cast<unsigned long long *>(w + own), (unsigned long long)flips);
}

// WIDE chains run 512-thread workgroups: the doubled lines formed in registers
// need more than the 128 VGPRs a 1024-thread workgroup leaves
template <bool WIDE>
__global__ __launch_bounds__(WIDE ? 512 : 1024) void fcc_bits_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const double *__restrict__ tables,
    const uint8_t *__restrict__ tri_g, const FccSched *__restrict__ sched,
    mcg_sample *__restrict__ series, int64_t series_cap,
    const uint8_t *__restrict__ active, int64_t n_passes, int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double tab[FCC_NTAB];
  __shared__ uint8_t tri[4096];
  __shared__ FccLists Ls;
  __shared__ FccSched S;
  __shared__ int item_start[FCC_SCHED_MAX + 1];
  __shared__ long long red[6][16];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  uint8_t *g_lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  uint64_t *w = reinterpret_cast<uint64_t *>(smem);
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < FCC_NTAB; i += nt) tab[i] = tables[P.tab_off + i];
  for (int i = t; i < 4096; i += nt) tri[i] = tri_g[i];
  const int L0 = P.L0, L1 = P.L1, L2 = P.L2;
  if (t < FCC_NCOL) fcc_fill_lists(Ls, t, L0, L1, L2);
  const int n_lines = 4 * L1 * L2;
  for (int ln = t; ln < n_lines; ln += nt) {
    uint64_t core = 0;
    for (int h = 0; h < L0; ++h) core |= (uint64_t)(g_lat[ln * L0 + h] & 1) << h;
    w[fcc_word_index(ln)] = WIDE ? core : (core | (core << L0));
  }
  __syncthreads();
  fcc_sched_setup<true>(sched, c, S, item_start, Ls, t);
  __syncthreads();
  int64_t pass = St->n_pass;
  int64_t N1 = St->M, P1 = St->S1, P2 = St->S2, P3 = St->S3, T3 = St->T3;
  int64_t nacc_tot = St->n_accept, n_samples = St->n_samples;
  const int wid = t >> 6, lane = t & 63, nw = nt >> 6;
  const int n_steps = S.n_steps;
  for (int64_t it = 0; it < n_passes; ++it) {
    FccDelta d;
    for (int s = 0; s < n_steps; ++s) {
      const int k0 = S.start[s], k1 = S.start[s + 1];
      for (int u = item_start[k0] + t; u < item_start[k1]; u += nt) {
        const int k = fcc_item_class(item_start, k0, k1, u);
        const FccClass kc = fcc_class(Ls, S.cls[k]);
        fcc_bits_unit<WIDE>(w, tab, tri, Ls, L0, L1, L2, kc, u - item_start[k], pass, P.key0,
                            P.key1, d);
      }
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
    for (int ww = 0; ww < nw; ++ww)
#pragma unroll
      for (int kk = 0; kk < 6; ++kk) a[kk] += red[kk][ww];
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
  for (int ln = t; ln < n_lines; ln += nt) {
    const uint64_t core = w[fcc_word_index(ln)];
    for (int h = 0; h < L0; ++h) g_lat[ln * L0 + h] = (uint8_t)((core >> h) & 1);
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

// ---------------------------------------------------------------------------
// Class-line kernel (lines of at most 32 sites, LDS-resident): one thread
// updates every site of the class being swept on one lattice line.  The
// line's 21 neighbourhood words are read once; each becomes one or two 32-bit
// vectors whose bit h is the neighbour of the line's site h at that offset
// (the doubled word makes every shift one alignbit), and carry-save adders sum
// the vectors position by position into bit planes of the counts: n2 (6
// inputs), n3 (24), the nearest-neighbour count n1 (12) and the triangle count
// t (the 24 products of nearest-neighbour pairs that are nearest neighbours of
// each other).  The class sites of the line are then decided one by one in
// slot order with the random words of the 4-site units of fcc_bits_unit
// (counter line*32 + q for slots 4q..4q+3) and the same FP64 products, so the
// trajectories equal the other FCC kernels'.  Flips go into the line with one
// store (no other thread touches it during the class step), and the count
// deltas come from population counts of the flip vector against the planes.

// full and half adders on 32 positions at once
__device__ __forceinline__ void fcc_fa(uint32_t a, uint32_t b, uint32_t c, uint32_t &s,
                                       uint32_t &cy) {
  const uint32_t t = a ^ b;
  s = t ^ c;
  cy = (t & c) | (a & ~t);
}
__device__ __forceinline__ void fcc_ha(uint32_t a, uint32_t b, uint32_t &s, uint32_t &cy) {
  s = a ^ b;
  cy = a & b;
}

// 6 one-bit vectors -> planes b[0..2]
__device__ __forceinline__ void fcc_count6(const uint32_t *x, uint32_t *b) {
  uint32_t s0, c0, s1, c1, c2;
  fcc_fa(x[0], x[1], x[2], s0, c0);
  fcc_fa(x[3], x[4], x[5], s1, c1);
  fcc_ha(s0, s1, b[0], c2);
  fcc_fa(c0, c1, c2, b[1], b[2]);
}

// 12 one-bit vectors -> planes b[0..3]
__device__ __forceinline__ void fcc_count12(const uint32_t *x, uint32_t *b) {
  uint32_t s[5], c[6], t0, t1, d[3];
  fcc_fa(x[0], x[1], x[2], s[0], c[0]);
  fcc_fa(x[3], x[4], x[5], s[1], c[1]);
  fcc_fa(x[6], x[7], x[8], s[2], c[2]);
  fcc_fa(x[9], x[10], x[11], s[3], c[3]);
  fcc_fa(s[0], s[1], s[2], s[4], c[4]);
  fcc_ha(s[4], s[3], b[0], c[5]);
  fcc_fa(c[0], c[1], c[2], t0, d[0]);
  fcc_fa(c[3], c[4], c[5], t1, d[1]);
  fcc_ha(t0, t1, b[1], d[2]);
  fcc_fa(d[0], d[1], d[2], b[2], b[3]);
}

// 24 one-bit vectors -> planes b[0..4]
__device__ __forceinline__ void fcc_count24(const uint32_t *x, uint32_t *b) {
  uint32_t s[8], c[12], u0, u1, u2, v[5], d[6], e0, e1, f[3];
#pragma unroll
  for (int i = 0; i < 8; ++i) fcc_fa(x[3 * i], x[3 * i + 1], x[3 * i + 2], s[i], c[i]);
  fcc_fa(s[0], s[1], s[2], u0, c[8]);
  fcc_fa(s[3], s[4], s[5], u1, c[9]);
  fcc_fa(u0, u1, s[6], u2, c[10]);
  fcc_ha(u2, s[7], b[0], c[11]);
#pragma unroll
  for (int i = 0; i < 4; ++i) fcc_fa(c[3 * i], c[3 * i + 1], c[3 * i + 2], v[i], d[i]);
  fcc_fa(v[0], v[1], v[2], v[4], d[4]);
  fcc_ha(v[4], v[3], b[1], d[5]);
  fcc_fa(d[0], d[1], d[2], e0, f[0]);
  fcc_fa(d[3], d[4], d[5], e1, f[1]);
  fcc_ha(e0, e1, b[2], f[2]);
  fcc_fa(f[0], f[1], f[2], b[3], b[4]);
}

// bit h of the result: bit (h + s) of the doubled line D (0 <= s <= 31)
__device__ __forceinline__ uint32_t fcc_vec(uint64_t D, uint32_t s) {
  return __builtin_amdgcn_alignbit((uint32_t)(D >> 32), (uint32_t)D, s);
}

__device__ __forceinline__ uint32_t fcc_bit(uint32_t v, int h) { return (v >> h) & 1u; }

// signed sum over the flipped sites of a count given by planes pl[0..nb-1]:
// sites 0 -> 1 (F0) add it, sites 1 -> 0 (F1) subtract it
template <int NB>
__device__ __forceinline__ int fcc_plane_delta(const uint32_t *pl, uint32_t F0, uint32_t F1) {
  int d = 0;
#pragma unroll
  for (int k = 0; k < NB; ++k)
    d += (__popc(pl[k] & F0) - __popc(pl[k] & F1)) << k;
  return d;
}

__device__ __forceinline__ void fcc_class_line(uint64_t *__restrict__ w,
                                               const double *__restrict__ tab,
                                               const FccLists &Ls, const uint32_t (*xmask)[2],
                                               int L0, int L1, int L2, int cx, int cy, int cz,
                                               int col, int b, int li0, int j, int g,
                                               int64_t pass, uint32_t key0, uint32_t key1,
                                               FccDelta &d) {
  const int py = b >> 1, pz = b & 1, p = py ^ pz;
  const int nyb = Ls.ny[cy][py];
  const int y = Ls.y[cy][py][li0 % nyb], z = Ls.z[cz][pz][li0 / nyb];
  const int ry = 2 * L1, rz = 2 * L2;
  int ys[5], zs[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    ys[j] = wrap_ring(y + j - 2, ry);
    zs[j] = wrap_ring(z + j - 2, rz) * ry;
  }
  auto line = [&](int dy, int dz) -> uint64_t {
    return w[fcc_word_index(ys[dy + 2] + zs[dz + 2])];
  };
  // the 21 words of the neighbourhood, all requested before any is used
  const int own = fcc_word_index(y + ry * z);
  const uint64_t Dn = w[own];
  const uint64_t S0 = line(-2, 0), S1 = line(2, 0), S2 = line(0, -2), S3 = line(0, 2);
  const uint64_t E0 = line(-1, -1), E1 = line(-1, 1), E2 = line(1, -1), E3 = line(1, 1);
  const uint64_t Q[8] = {line(-2, -1), line(-2, 1), line(2, -1), line(2, 1),
                         line(-1, -2), line(1, -2), line(-1, 2), line(1, 2)};
  const uint64_t N[4] = {line(-1, 0), line(1, 0), line(0, -1), line(0, 1)};
  // the random words meanwhile: slots 4q..4q+3 of the class sites take the
  // words of call q (at most 16 sites, four calls, when L0 <= 32); of the g
  // lanes that share the line, lane j takes the calls q = j, j + g, ...
  const uint32_t M0 = xmask[cx][p];
  const int nq = (__popc(M0) + 3) >> 2;
  const uint32_t ctr = (uint32_t)(y + ry * z) * 32u;
  u32x4 r[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (q < nq && (q & (g - 1)) == j)
      r[q] = philox4x32_10(ctr + (uint32_t)q, (uint32_t)pass, (uint32_t)(pass >> 32),
                           (uint32_t)col, key0, key1);
  const uint32_t sm = (uint32_t)(L0 - 1);       // shift giving the neighbour at h - 1
  const uint32_t sA = p ? 0u : sm, sB = (uint32_t)p;  // odd lines: sites h+p-1, h+p
  const uint32_t O0 = (uint32_t)Dn;
  // second shell: x -+ 2 on the own line and (0, +-2, 0), (0, 0, +-2)
  uint32_t in2[6] = {fcc_vec(Dn, sm), fcc_vec(Dn, 1u), (uint32_t)S0, (uint32_t)S1, (uint32_t)S2,
                     (uint32_t)S3};
  uint32_t nn[12], in3[24];
  nn[8] = (uint32_t)E0;
  nn[9] = (uint32_t)E1;
  nn[10] = (uint32_t)E2;
  nn[11] = (uint32_t)E3;
  in3[0] = fcc_vec(E0, sm);
  in3[1] = fcc_vec(E0, 1u);
  in3[2] = fcc_vec(E1, sm);
  in3[3] = fcc_vec(E1, 1u);
  in3[4] = fcc_vec(E2, sm);
  in3[5] = fcc_vec(E2, 1u);
  in3[6] = fcc_vec(E3, sm);
  in3[7] = fcc_vec(E3, 1u);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    in3[8 + 2 * i] = fcc_vec(Q[i], sA);
    in3[9 + 2 * i] = fcc_vec(Q[i], sB);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    nn[2 * i] = fcc_vec(N[i], sA);
    nn[2 * i + 1] = fcc_vec(N[i], sB);
  }
  uint32_t n2p[3], n3p[5], n1p[4], tp[5], tri[24];
  fcc_count6(in2, n2p);
  fcc_count24(in3, n3p);
  fcc_count12(nn, n1p);
#pragma unroll
  for (int i = 0; i < 24; ++i) tri[i] = nn[FCC_TRI_PAIRS[i][0]] & nn[FCC_TRI_PAIRS[i][1]];
  fcc_count24(tri, tp);
  // the class sites four at a time: positions, table indices (each count bit
  // extracted and weighted by its place in the index), the eight table reads
  // in flight together, then the decisions
  uint32_t M = M0, F = 0u;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q < nq && (q & (g - 1)) == j) {
      int hh[4], iA[4], iB[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int h = M != 0u ? __builtin_ctz(M) : 0;
        M &= M - 1u;
        hh[v] = h;
        auto bit = [&](uint32_t pl) { return __builtin_amdgcn_ubfe(pl, (uint32_t)h, 1u); };
        const uint32_t sg = bit(O0);
        const uint32_t n1 = bit(n1p[3]) << 3 | bit(n1p[2]) << 2 | bit(n1p[1]) << 1 | bit(n1p[0]);
        const uint32_t n2 = bit(n2p[2]) << 2 | bit(n2p[1]) << 1 | bit(n2p[0]);
        const uint32_t n3 = bit(n3p[4]) << 4 | bit(n3p[3]) << 3 | bit(n3p[2]) << 2 |
                            bit(n3p[1]) << 1 | bit(n3p[0]);
        const uint32_t t = bit(tp[4]) << 4 | bit(tp[3]) << 3 | bit(tp[2]) << 2 |
                           bit(tp[1]) << 1 | bit(tp[0]);
        iA[v] = (int)(sg * FCC_TAB_A + n1 * (FCC_Z2 + 1) + n2);
        iB[v] = (int)(2 * FCC_TAB_A + sg * FCC_TAB_B + n3 * (FCC_ZT + 1) + t);
      }
      double A[4], B[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        A[v] = tab[iA[v]];
        B[v] = tab[iB[v]];
      }
      const int left = __popc(M0) - 4 * q;  // class sites from slot 4q on
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double uu = ((double)r[q].v[v] + 0.5) * 0x1p-32;
        if (v < left && uu < A[v] * B[v]) F |= 1u << hh[v];
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) M &= M - 1u;  // another lane's slots
    }
  }
  if (F != 0u) {
    // the g lanes of a line flip disjoint sites; every read of the line in
    // this class step came before (same wavefront, earlier instruction)
    atomicXor(reinterpret_cast<unsigned long long *>(w + own),
              (unsigned long long)((uint64_t)F | ((uint64_t)F << L0)));
    const uint32_t F1 = F & O0, F0 = F & ~O0;
    d.N1 += __popc(F0) - __popc(F1);
    d.P1 += fcc_plane_delta<4>(n1p, F0, F1);
    d.P2 += fcc_plane_delta<3>(n2p, F0, F1);
    d.P3 += fcc_plane_delta<5>(n3p, F0, F1);
    d.T += fcc_plane_delta<5>(tp, F0, F1);
    d.acc += __popc(F);
  }
}

__global__ __launch_bounds__(512) void fcc_lines_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const double *__restrict__ tables,
    const FccSched *__restrict__ sched, mcg_sample *__restrict__ series, int64_t series_cap,
    const uint8_t *__restrict__ active, int64_t n_passes, int64_t sample_period) {
  const int c = blockIdx.x;
  if (!active[c]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double tab[FCC_NTAB];
  __shared__ FccLists Ls;
  __shared__ FccSched S;
  __shared__ int item_start[FCC_SCHED_MAX + 1];
  __shared__ uint32_t xmask[FCC_NCOL][2];
  __shared__ long long red[6][8];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  uint8_t *g_lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  uint64_t *w = reinterpret_cast<uint64_t *>(smem);
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < FCC_NTAB; i += nt) tab[i] = tables[P.tab_off + i];
  const int L0 = P.L0, L1 = P.L1, L2 = P.L2;
  if (t < FCC_NCOL) fcc_fill_lists(Ls, t, L0, L1, L2);
  const int n_lines = 4 * L1 * L2;
  for (int ln = t; ln < n_lines; ln += nt) {
    uint64_t core = 0;
    for (int h = 0; h < L0; ++h) core |= (uint64_t)(g_lat[ln * L0 + h] & 1) << h;
    w[fcc_word_index(ln)] = core | (core << L0);
  }
  __syncthreads();
  if (t < 2 * FCC_NCOL) {  // positions of colour t>>1 on lines of x parity t&1
    const int a = t >> 1, pr = t & 1;
    uint32_t m = 0u;
    for (int s = 0; s < Ls.nx[a][pr]; ++s) m |= 1u << Ls.xh[a][pr][s];
    xmask[a][pr] = m;
  }
  fcc_sched_setup<false>(sched, c, S, item_start, Ls, t);
  __syncthreads();
  int64_t pass = St->n_pass;
  int64_t N1 = St->M, P1 = St->S1, P2 = St->S2, P3 = St->S3, T3 = St->T3;
  int64_t nacc_tot = St->n_accept, n_samples = St->n_samples;
  const int wid = t >> 6, lane = t & 63, nw = nt >> 6;
  const int n_steps = S.n_steps;
  for (int64_t it = 0; it < n_passes; ++it) {
    FccDelta d;
    for (int s = 0; s < n_steps; ++s) {
      const int k0 = S.start[s], k1 = S.start[s + 1];
      const int i0 = item_start[k0], n_items = item_start[k1] - i0;
      // lanes per line: 4 or 2 when the step's lines leave threads over
      // (shorter steps for small classes), else one
      const int g = 4 * n_items <= nt ? 4 : 2 * n_items <= nt ? 2 : 1;
      for (int li = t / g; li < n_items; li += nt / g) {
        const int k = fcc_item_class(item_start, k0, k1, i0 + li);
        const int col = S.cls[k];
        const int cx = col % FCC_NCOL, cy = (col / FCC_NCOL) % FCC_NCOL,
                  cz = col / (FCC_NCOL * FCC_NCOL);
        int lstart[4];
        lstart[0] = 0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const int py = b >> 1, pz = b & 1;
          const int nl = Ls.nx[cx][py ^ pz] > 0 ? Ls.ny[cy][py] * Ls.nz[cz][pz] : 0;
          lstart[b + 1] = lstart[b] + nl;
        }
        const int lc = i0 + li - item_start[k];  // line within its class
        const int b = (lc >= lstart[1]) + (lc >= lstart[2]) + (lc >= lstart[3]);
        const int li0 = lc - lstart[b];
        fcc_class_line(w, tab, Ls, xmask, L0, L1, L2, cx, cy, cz, col, b, li0, t % g, g, pass,
                       P.key0, P.key1, d);
      }
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
    for (int ww = 0; ww < nw; ++ww)
#pragma unroll
      for (int kk = 0; kk < 6; ++kk) a[kk] += red[kk][ww];
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
  for (int ln = t; ln < n_lines; ln += nt) {
    const uint64_t core = w[fcc_word_index(ln)];
    for (int h = 0; h < L0; ++h) g_lat[ln * L0 + h] = (uint8_t)((core >> h) & 1);
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

// The triangle pairs of fcc_lines_kernel against a triangle table in the
// same bit order: true when they give the table's count for every mask.
bool fcc_tri_pairs_check(const std::vector<uint8_t> &tri_lines) {
  if (tri_lines.size() != 4096) return false;
  for (int m = 0; m < 4096; ++m) {
    int t = 0;
    for (auto const &pr : FCC_TRI_PAIRS) t += ((m >> pr[0]) & 1) & ((m >> pr[1]) & 1);
    if (t != tri_lines[m]) return false;
  }
  return true;
}

void launch_fcc(const CheckerParams *P, CheckerState *St, int8_t *lat_pool, const double *tables,
                const uint8_t *tri, const uint8_t *tri_lines, const FccSched *sched,
                mcg_sample *series, int64_t series_cap, const uint8_t *active, int n_chains,
                int64_t n_passes, int64_t sample_period, int lds_bytes, bool lattice_in_lds,
                bool bit_lines, bool wide_lines, bool class_lines, int block,
                hipStream_t stream) {
  if (bit_lines && !wide_lines && class_lines)
    hipLaunchKernelGGL(fcc_lines_kernel, dim3(n_chains), dim3(block > 512 ? 512 : block),
                       lds_bytes, stream, P, St, lat_pool, tables, sched, series, series_cap,
                       active, n_passes, sample_period);
  else if (bit_lines && wide_lines)
    hipLaunchKernelGGL(fcc_bits_kernel<true>, dim3(n_chains), dim3(block), lds_bytes, stream, P,
                       St, lat_pool, tables, tri_lines, sched, series, series_cap, active,
                       n_passes, sample_period);
  else if (bit_lines)
    hipLaunchKernelGGL(fcc_bits_kernel<false>, dim3(n_chains), dim3(block), lds_bytes, stream, P,
                       St, lat_pool, tables, tri_lines, sched, series, series_cap, active,
                       n_passes, sample_period);
  else if (lattice_in_lds)
    hipLaunchKernelGGL(fcc_kernel<true>, dim3(n_chains), dim3(block), lds_bytes, stream, P, St,
                       lat_pool, tables, tri, sched, series, series_cap, active, n_passes,
                       sample_period);
  else
    hipLaunchKernelGGL(fcc_kernel<false>, dim3(n_chains), dim3(block), 0, stream, P, St,
                       lat_pool, tables, tri, sched, series, series_cap, active, n_passes,
                       sample_period);
}

hipError_t fcc_set_max_lds(int bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(fcc_bits_kernel<false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void *>(fcc_lines_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void *>(fcc_bits_kernel<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(fcc_kernel<true>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
