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
// Device layout ("split rows"): every lattice line (fixed j,k) of L0 sites is
// stored as its L0/2 even-i sites followed by its L0/2 odd-i sites, one byte
// per site holding 1 for spin +1 and 0 for -1.  A colour owns one half of each
// of its lines, so a sweep reads and writes contiguous runs: the site at half
// a, index h (i = a + 2h) has its in-line neighbours in the other half at
// h-1+a and h+a, and its j/k neighbours at index h of the same or the other
// half of the neighbouring lines.
//
// Two code paths give identical results:
//   chunk path (L0 % 32 == 0): a thread owns 16 consecutive sites of one half
//     line; neighbour counts are formed four sites per 32-bit add (byte lanes)
//     from 16-byte loads, the in-line shifts with v_alignbyte;
//   quad path (any even L0): a thread owns 4 consecutive sites, byte loads.
// Random numbers: site h of colour-class line ln (lines of one colour in
// order) draws a 32-bit uniform from Philox4x32-10 keyed by the chain seed:
// its high 16 bits from call (ln*ceil(L0/16) + h/8, pass lo, pass hi, colour),
// its low 16 bits, read only when they can decide, from call
// (ln*L0/2 + h, pass lo, pass hi, colour | 2^31) (see SiteRng).  A site's
// stream depends only on (seed, pass, colour, site): both paths, the
// LDS-resident kernel, the HBM kernels, a chain alone and a chain in a batch
// agree bit for bit.
//
// Acceptance: class (spin, up-neighbour counts) -> 32-bit threshold t with
// accept iff r <= t (t = 2^32-1 when dE <= 0), tabulated per chain on the host
// with exp(-beta*dE) at index checker_class(b, c1, c2); no exp per attempt.  Observables are exact integers
// (S1 = sum_NN s s, S2 = sum_NNN s s, M = sum s) updated from per-flip deltas
// reduced per workgroup.
#include "internal.hh"

namespace mcg {

struct LineGeom {
  int32_t half;     // sites per half line (L0/2)
  int32_t qpl;      // quads per half line, ceil(half/4)
  int32_t opl;      // 8-site groups per half line, ceil(half/8) (Philox calls)
  int32_t nl;       // lines per colour
  int32_t chunked;  // 16-site chunk path (half % 16 == 0)
};

__device__ __forceinline__ LineGeom line_geom(const CheckerParams &P) {
  LineGeom g;
  g.half = P.L0 >> 1;
  g.qpl = (g.half + 3) >> 2;
  g.opl = (g.half + 7) >> 3;
  g.chunked = ((g.half & 15) == 0) && !P.quad_only;
  if (P.ncolor == 2)
    g.nl = P.L1 * P.L2;
  else if (P.ncolor == 4)
    g.nl = P.L1 >> 1;
  else
    g.nl = (P.L1 >> 1) * (P.L2 >> 1);
  return g;
}

__device__ __forceinline__ int wrapm(int a, int L) { return a < 0 ? a + L : (a >= L ? a - L : a); }

// Line (j, k) and own half a of colour-class line ln.
__device__ __forceinline__ void line_of(const CheckerParams &P, int color, int ln, int &j, int &k,
                                        int &a) {
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
}

// Base pointers of the half lines around line (j, k): own line, j+-1, and in
// 3D k+-1 and the four (j+-1, k+-1) lines; `a` selects the half.
struct Lines {
  uint8_t *own;
  const uint8_t *jp, *jm, *kp, *km, *jpkp, *jpkm, *jmkp, *jmkm;
};

__device__ __forceinline__ Lines lines_of(const CheckerParams &P, uint8_t *lat, int j, int k) {
  const int64_t L0 = P.L0, plane = (int64_t)P.L0 * P.L1;
  const int jp = wrapm(j + 1, P.L1), jm = wrapm(j - 1, P.L1);
  Lines Ls;
  Ls.own = lat + (int64_t)k * plane + j * L0;
  Ls.jp = lat + (int64_t)k * plane + jp * L0;
  Ls.jm = lat + (int64_t)k * plane + jm * L0;
  Ls.kp = Ls.km = Ls.jpkp = Ls.jpkm = Ls.jmkp = Ls.jmkm = Ls.own;
  if (P.ndim == 3) {
    const int kp = wrapm(k + 1, P.L2), km = wrapm(k - 1, P.L2);
    Ls.kp = lat + (int64_t)kp * plane + j * L0;
    Ls.km = lat + (int64_t)km * plane + j * L0;
    Ls.jpkp = lat + (int64_t)kp * plane + jp * L0;
    Ls.jpkm = lat + (int64_t)km * plane + jp * L0;
    Ls.jmkp = lat + (int64_t)kp * plane + jm * L0;
    Ls.jmkm = lat + (int64_t)km * plane + jm * L0;
  }
  return Ls;
}

// Random words.  A site's 32-bit uniform r is split in two 16-bit halves: the
// high half comes from the main stream, one Philox4x32-10 call per 8 sites of
// a colour-class half line (counter (ln*opl + h/8, pass lo, pass hi, colour);
// site h takes bits 16*(h&1) of word (h&7)>>1), and the low half, needed only
// when the high half equals the threshold's (probability 2^-16), from word 0
// of the refinement call (ln*half + h, pass lo, pass hi, colour | 2^31).  The
// acceptance r <= t is therefore decided exactly as for a 32-bit uniform,
// with half the Philox work of one call per 4 sites.
struct SiteRng {
  uint32_t key0, key1, pass_lo, pass_hi, color;
};

__device__ __forceinline__ uint32_t hi16_of(const u32x4 &r, int e) {
  const uint32_t w = r.v[(e & 7) >> 1];
  return (e & 1) ? (w >> 16) : (w & 0xFFFFu);
}

__device__ __forceinline__ uint32_t refine_lo16(const SiteRng &rc, uint32_t site_ctr) {
  return philox4x32_10(site_ctr, rc.pass_lo, rc.pass_hi, rc.color | 0x80000000u, rc.key0,
                       rc.key1).v[0] & 0xFFFFu;
}

// Acceptance class of (old spin bit b, up-neighbour counts c1, c2) and the
// decision r <= t from the high half when it settles it: returns 1 (flip),
// 0 (keep) or -1 (the low half decides: flip iff lo16 <= (t & 0xFFFF)).
__device__ __forceinline__ int decide_hi(int b, int c1, int c2, uint32_t hi16,
                                         const uint32_t *__restrict__ tab, int z1, int z2,
                                         uint32_t &t) {
  const int cls = checker_class(b, c1, c2);
  t = tab[cls];
  if (t == 0xFFFFFFFFu) return 1;
  const uint32_t thi = t >> 16;
  return hi16 < thi ? 1 : (hi16 > thi ? 0 : -1);
}

__device__ __forceinline__ void count_flip(int b, int c1, int c2, int z1, int z2, int &dS1,
                                           int &dS2, int &dM, int &nacc) {
  const int s = 2 * b - 1;
  dS1 += -2 * s * (2 * c1 - z1);
  dS2 += -2 * s * (2 * c2 - z2);
  dM += -2 * s;
  nacc += 1;
}

// --- quad path ------------------------------------------------------------------
__device__ __forceinline__ void update_quad(const CheckerParams &P, const LineGeom &g,
                                            uint8_t *__restrict__ lat,
                                            const uint32_t *__restrict__ tab, int color,
                                            int64_t pass, int64_t qd, int &dS1, int &dS2,
                                            int &dM, int &nacc) {
  const int ln = (int)(qd / g.qpl);
  const int qi = (int)(qd - (int64_t)ln * g.qpl);
  int j, k, a;
  line_of(P, color, ln, j, k, a);
  const Lines Ls = lines_of(P, lat, j, k);
  const int half = g.half;
  const int oa = a * half, ob = (1 - a) * half;  // own half, other half offsets
  const SiteRng rc{P.key0, P.key1, (uint32_t)pass, (uint32_t)(pass >> 32), (uint32_t)color};
  const u32x4 r = philox4x32_10((uint32_t)((int64_t)ln * g.opl + (qi >> 1)), rc.pass_lo,
                                rc.pass_hi, rc.color, P.key0, P.key1);
  const int z1 = P.z1, z2 = P.z2;
  uint32_t need = 0u;  // sites whose low random half decides (threshold kept in tneed)
  uint32_t tneed[4] = {0u, 0u, 0u, 0u};
  int cneed[4] = {0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int h = 4 * qi + u;
    if (h >= half) break;
    const int he = wrapm(h + a, half), hw = wrapm(h - 1 + a, half);
    const int b = Ls.own[oa + h];
    int c1 = Ls.own[ob + he] + Ls.own[ob + hw] + Ls.jp[oa + h] + Ls.jm[oa + h];
    int c2 = 0;
    if (P.ndim == 3) c1 += Ls.kp[oa + h] + Ls.km[oa + h];
    if (z2 > 0) {
      c2 = Ls.jp[ob + he] + Ls.jp[ob + hw] + Ls.jm[ob + he] + Ls.jm[ob + hw];
      if (P.ndim == 3)
        c2 += Ls.kp[ob + he] + Ls.kp[ob + hw] + Ls.km[ob + he] + Ls.km[ob + hw] +
              Ls.jpkp[oa + h] + Ls.jpkm[oa + h] + Ls.jmkp[oa + h] + Ls.jmkm[oa + h];
    }
    uint32_t t;
    const int flip = decide_hi(b, c1, c2, hi16_of(r, h), tab, z1, z2, t);
    if (flip > 0) {
      count_flip(b, c1, c2, z1, z2, dS1, dS2, dM, nacc);
      Ls.own[oa + h] = (uint8_t)(b ^ 1);
    } else if (flip < 0) {
      need |= 1u << u;
      tneed[u] = t;
      cneed[u] = (b << 16) | (c1 << 8) | c2;
    }
  }
  while (need) {  // rare: ties of the high halves, decided by the refinement words
    const int u = __builtin_ctz(need);
    need &= need - 1u;
    const int h = 4 * qi + u;
    const uint32_t t = u == 0 ? tneed[0] : u == 1 ? tneed[1] : u == 2 ? tneed[2] : tneed[3];
    const int cc = u == 0 ? cneed[0] : u == 1 ? cneed[1] : u == 2 ? cneed[2] : cneed[3];
    if (refine_lo16(rc, (uint32_t)((int64_t)ln * half + h)) <= (t & 0xFFFFu)) {
      const int b = cc >> 16;
      count_flip(b, (cc >> 8) & 0xFF, cc & 0xFF, z1, z2, dS1, dS2, dM, nacc);
      Ls.own[oa + h] = (uint8_t)(b ^ 1);
    }
  }
}

// --- chunk path (16 sites, SWAR) -------------------------------------------------
__device__ __forceinline__ uint4 ld16(const uint8_t *p) {
  return *reinterpret_cast<const uint4 *>(p);
}

// element e <- e+1, `last` (low byte) enters at element 15
__device__ __forceinline__ uint4 shift_down(uint4 v, uint32_t last) {
  uint4 r;
  r.x = __builtin_amdgcn_alignbyte(v.y, v.x, 1);
  r.y = __builtin_amdgcn_alignbyte(v.z, v.y, 1);
  r.z = __builtin_amdgcn_alignbyte(v.w, v.z, 1);
  r.w = __builtin_amdgcn_alignbyte(last, v.w, 1);
  return r;
}

// element e <- e-1, `first` (low byte) enters at element 0
__device__ __forceinline__ uint4 shift_up(uint4 v, uint32_t first) {
  uint4 r;
  r.x = __builtin_amdgcn_alignbyte(v.x, first << 24, 3);
  r.y = __builtin_amdgcn_alignbyte(v.y, v.x, 3);
  r.z = __builtin_amdgcn_alignbyte(v.z, v.y, 3);
  r.w = __builtin_amdgcn_alignbyte(v.w, v.z, 3);
  return r;
}

__device__ __forceinline__ uint4 add4(uint4 a, uint4 b) {
  return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// In-line neighbour pair of the 16 own sites [h0, h0+16) from the other half
// line `oth` (base of that half): returns east + west.
__device__ __forceinline__ uint4 inline_pair(const uint8_t *oth, int h0, int a, int half) {
  const uint4 v = ld16(oth + h0);
  if (a == 0) {  // east = h, west = h-1
    const uint32_t first = oth[wrapm(h0 - 1, half)];
    return add4(v, shift_up(v, first));
  }
  // a == 1: west = h, east = h+1
  const uint32_t last = oth[wrapm(h0 + 16, half)];
  return add4(v, shift_down(v, last));
}

__device__ __forceinline__ uint32_t word_of(uint4 v, int w) {
  return w == 0 ? v.x : (w == 1 ? v.y : (w == 2 ? v.z : v.w));
}

// The 16 sites [h0, h0+16) of half a of the line whose half lines are at Ls,
// colour-class line ln (main-stream Philox counters ln*opl + h0/8 and the next).
__device__ __forceinline__ void update_chunk_at(const CheckerParams &P, const LineGeom &g,
                                                const Lines &Ls, int a, int ln, int h0,
                                                const uint32_t *__restrict__ tab, int color,
                                                int64_t pass, int &dS1, int &dS2, int &dM,
                                                int &nacc) {
  const int half = g.half;
  const int oa = a * half, ob = (1 - a) * half;
  const int z1 = P.z1, z2 = P.z2;

  uint4 own = ld16(Ls.own + oa + h0);
  uint4 c1 = add4(inline_pair(Ls.own + ob, h0, a, half),
                  add4(ld16(Ls.jp + oa + h0), ld16(Ls.jm + oa + h0)));
  if (P.ndim == 3) c1 = add4(c1, add4(ld16(Ls.kp + oa + h0), ld16(Ls.km + oa + h0)));
  uint4 c2 = make_uint4(0, 0, 0, 0);
  if (z2 > 0) {
    c2 = add4(inline_pair(Ls.jp + ob, h0, a, half), inline_pair(Ls.jm + ob, h0, a, half));
    if (P.ndim == 3) {
      c2 = add4(c2, add4(inline_pair(Ls.kp + ob, h0, a, half),
                         inline_pair(Ls.km + ob, h0, a, half)));
      c2 = add4(c2, add4(add4(ld16(Ls.jpkp + oa + h0), ld16(Ls.jpkm + oa + h0)),
                         add4(ld16(Ls.jmkp + oa + h0), ld16(Ls.jmkm + oa + h0))));
    }
  }
  const SiteRng rc{P.key0, P.key1, (uint32_t)pass, (uint32_t)(pass >> 32), (uint32_t)color};
  const uint32_t o0 = (uint32_t)((int64_t)ln * g.opl + (h0 >> 3));
  // class byte of every site, checker_class(b, c1, c2) (< 256: no carries
  // between the bytes of a word)
  uint32_t fw[4] = {0u, 0u, 0u, 0u};  // 1 in the byte of every site that flips
  uint32_t need = 0u;                 // sites whose low random half decides
#pragma unroll
  for (int hw = 0; hw < 2; ++hw) {
    const u32x4 r = philox4x32_10(o0 + hw, rc.pass_lo, rc.pass_hi, rc.color, P.key0, P.key1);
#pragma unroll
    for (int w = 2 * hw; w < 2 * hw + 2; ++w) {
      const uint32_t cw = (word_of(own, w) << 7) + (word_of(c1, w) << 4) + word_of(c2, w);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t thi = tab[(cw >> (8 * u)) & 0xFFu] >> 16;
        const uint32_t hi = hi16_of(r, 4 * w + u);
        fw[w] |= (hi < thi ? 1u : 0u) << (8 * u);
        need |= (hi == thi ? 1u : 0u) << (4 * w + u);
      }
    }
  }
  while (need) {  // rare: ties of the high halves, decided by the refinement words
    const int e = __builtin_ctz(need);
    need &= need - 1u;
    const int w = e >> 2, u = e & 3;
    const uint32_t cw = (word_of(own, w) << 7) + (word_of(c1, w) << 4) + word_of(c2, w);
    const uint32_t t = tab[(cw >> (8 * u)) & 0xFFu];
    if (refine_lo16(rc, (uint32_t)((int64_t)ln * half + h0 + e)) <= (t & 0xFFFFu)) {
      if (w == 0) fw[0] |= 1u << (8 * u);
      if (w == 1) fw[1] |= 1u << (8 * u);
      if (w == 2) fw[2] |= 1u << (8 * u);
      if (w == 3) fw[3] |= 1u << (8 * u);
    }
  }
  // observable changes, four sites per word: a flip of a site with spin bit b
  // and up-neighbour counts c1, c2 changes S1 by -2(2b-1)(2c1-z1), S2 by
  // -2(2b-1)(2c2-z2) and M by -2(2b-1); byte dot products give the sums over
  // the flipped sites of 1, b, c1, b*c1, c2 and b*c2
  const uint32_t ones = 0x01010101u;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t f = fw[w];
    if (!f) continue;
    const uint32_t fb = f & word_of(own, w);
    const int nf = (int)__builtin_amdgcn_udot4(f, ones, 0u, false);
    const int nb = (int)__builtin_amdgcn_udot4(fb, ones, 0u, false);
    const int s1 = (int)__builtin_amdgcn_udot4(f, word_of(c1, w), 0u, false);
    const int b1 = (int)__builtin_amdgcn_udot4(fb, word_of(c1, w), 0u, false);
    const int s2 = (int)__builtin_amdgcn_udot4(f, word_of(c2, w), 0u, false);
    const int b2 = (int)__builtin_amdgcn_udot4(fb, word_of(c2, w), 0u, false);
    dM += -2 * (2 * nb - nf);
    dS1 += -2 * (4 * b1 - 2 * z1 * nb - 2 * s1 + z1 * nf);
    dS2 += -2 * (4 * b2 - 2 * z2 * nb - 2 * s2 + z2 * nf);
    nacc += nf;
  }
  if (fw[0] | fw[1] | fw[2] | fw[3]) {
    own.x ^= fw[0];
    own.y ^= fw[1];
    own.z ^= fw[2];
    own.w ^= fw[3];
    *reinterpret_cast<uint4 *>(Ls.own + oa + h0) = own;
  }
}

__device__ __forceinline__ void update_chunk(const CheckerParams &P, const LineGeom &g,
                                             uint8_t *__restrict__ lat,
                                             const uint32_t *__restrict__ tab, int color,
                                             int64_t pass, int64_t cd, int &dS1, int &dS2,
                                             int &dM, int &nacc) {
  const int cpl = g.half >> 4;  // chunks per half line
  const int ln = (int)(cd / cpl);
  const int h0 = (int)(cd - (int64_t)ln * cpl) << 4;
  int j, k, a;
  line_of(P, color, ln, j, k, a);
  const Lines Ls = lines_of(P, lat, j, k);
  update_chunk_at(P, g, Ls, a, ln, h0, tab, color, pass, dS1, dS2, dM, nacc);
}

// Work units of one colour of one chain: chunks (chunk path) or quads.
__device__ __forceinline__ bool chunk_path(const LineGeom &g) { return g.chunked != 0; }

__device__ __forceinline__ int64_t n_units(const LineGeom &g) {
  return (int64_t)g.nl * (chunk_path(g) ? (g.half >> 4) : g.qpl);
}

__device__ __forceinline__ void update_unit(const CheckerParams &P, const LineGeom &g,
                                            uint8_t *lat, const uint32_t *tab, int color,
                                            int64_t pass, int64_t unit, int &dS1, int &dS2,
                                            int &dM, int &nacc) {
  if (chunk_path(g))
    update_chunk(P, g, lat, tab, color, pass, unit, dS1, dS2, dM, nacc);
  else
    update_quad(P, g, lat, tab, color, pass, unit, dS1, dS2, dM, nacc);
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
  __shared__ int red[4][16];
  __shared__ uint32_t tab[CHECKER_NTAB];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  uint8_t *g_lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  uint8_t *lat = reinterpret_cast<uint8_t *>(smem);
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < CHECKER_NTAB; i += nt) tab[i] = tables[P.tab_off + i];
  const int64_t n = P.n;
  const int64_t n16 = (n + 15) >> 4;  // lattices are padded to 16 bytes
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(g_lat);
    uint4 *dst = reinterpret_cast<uint4 *>(lat);
    for (int64_t i = t; i < n16; i += nt) dst[i] = src[i];
  }
  __syncthreads();
  const LineGeom g = line_geom(P);
  const int64_t nu = n_units(g);
  int64_t pass = St->n_pass;
  int64_t S1 = St->S1, S2 = St->S2, M = St->M, nacc_tot = St->n_accept;
  int64_t n_samples = St->n_samples;
  const int wid = t >> 6, lane = t & 63, nw = nt >> 6;
  for (int64_t p = 0; p < n_passes; ++p) {
    int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
    for (int color = 0; color < P.ncolor; ++color) {
      for (int64_t u = t; u < nu; u += nt)
        update_unit(P, g, lat, tab, color, pass, u, dS1, dS2, dM, nacc);
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
        mcg_sample rec = {};
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
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(lat);
    uint4 *dst = reinterpret_cast<uint4 *>(g_lat);
    for (int64_t i = t; i < n16; i += nt) dst[i] = src[i];
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

hipError_t checker_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(checker_lds_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// ---------------------------------------------------------------------------
// HBM work-list kernel: one launch per colour over a list of (chain, item)
// entries; an item is UNITS_PER_ITEM work units of one chain.
// ---------------------------------------------------------------------------
constexpr int HBM_BLOCK = 256;
constexpr int MAX_UNITS_PER_THREAD = 4;
constexpr int UNITS_PER_ITEM = HBM_BLOCK * MAX_UNITS_PER_THREAD;

// A work item is UNITS_PER_ITEM units of one chain; a block covers
// HBM_BLOCK * UPT units, so an item takes MAX_UNITS_PER_THREAD / UPT blocks.
template <int UPT>
__global__ __launch_bounds__(HBM_BLOCK) void checker_color_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ work, int color) {
  constexpr int BLOCKS_PER_ITEM = MAX_UNITS_PER_THREAD / UPT;
  __shared__ int red[4][HBM_BLOCK / 64];
  const int item = blockIdx.x / BLOCKS_PER_ITEM;
  const int sub = blockIdx.x - item * BLOCKS_PER_ITEM;
  const int c = work[2 * item];
  const int64_t u0 =
      (int64_t)work[2 * item + 1] * UNITS_PER_ITEM + (int64_t)sub * HBM_BLOCK * UPT;
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  const LineGeom g = line_geom(P);
  const int64_t nu = n_units(g);
  const int64_t pass = St->n_pass;
  const uint32_t *tab = tables + P.tab_off;
  uint8_t *lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < UPT; ++u) {
    const int64_t unit = u0 + (int64_t)u * HBM_BLOCK + t;
    if (unit < nu) update_unit(P, g, lat, tab, color, pass, unit, dS1, dS2, dM, nacc);
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

// ---------------------------------------------------------------------------
// HBM line-class kernel (4 and 8 colourings): the two colours of one line
// class (halves a = 0 and a = 1 of the same lines) in one launch.  Colour
// 2m+1 depends on colour 2m only through the in-line neighbours of the same
// line (its other neighbours belong to other line classes), so a workgroup that
// owns whole lines updates half 0, synchronises, and updates half 1: half the
// launches, and half the passes over the lattice, of one launch per colour.
// ---------------------------------------------------------------------------
constexpr int PAIR_BLOCK = 256;

// Work items are (chain, first colour-class line, lines, 0): lines straight
// from HBM (every class line re-reads its neighbour lines).
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup i to XCD
// i % 8), and each XCD has its own L2.  Work items are ordered so that nearby
// items read the same neighbour lines; this bijection hands every XCD one
// contiguous run of items (XCD x takes items [start_x, start_x + count_x) in
// order), so those shared lines meet in one L2.  Results do not depend on the
// order: items update disjoint lines and sum integers.
__device__ __forceinline__ int xcd_item(int i, int n) {
  const int x = i & 7, local = i >> 3, q = n >> 3, r = n & 7;
  return x < r ? x * (q + 1) + local : r * (q + 1) + (x - r) * q + local;
}

__global__ __launch_bounds__(PAIR_BLOCK) void checker_pair_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ work, int line_class) {
  __shared__ int red[4][PAIR_BLOCK / 64];
  const int32_t *item = work + 4 * xcd_item(blockIdx.x, gridDim.x);
  const int c = item[0];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  const LineGeom g = line_geom(P);
  const int64_t pass = St->n_pass;
  const uint32_t *tab = tables + P.tab_off;
  uint8_t *lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  const int upl = chunk_path(g) ? (g.half >> 4) : g.qpl;  // units per half line
  int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
  const int l0 = item[1], nlines = item[2];
  // Both colours of a group of lines (one unit per thread) before the next
  // group: colour 1 of a line reads only that line's colour-0 half among the
  // sites this launch changes, so the order is free, and a group's lines are
  // still in cache for its second colour.
  const int per_group = upl < PAIR_BLOCK ? PAIR_BLOCK / upl : 1;
  for (int g0 = 0; g0 < nlines; g0 += per_group) {
    const int gl = nlines - g0 < per_group ? nlines - g0 : per_group;
    for (int a = 0; a < 2; ++a) {
      const int color = 2 * line_class + a;
      for (int task = threadIdx.x; task < gl * upl; task += PAIR_BLOCK) {
        const int ln = l0 + g0 + task / upl;
        const int64_t unit = (int64_t)ln * upl + (task - (task / upl) * upl);
        update_unit(P, g, lat, tab, color, pass, unit, dS1, dS2, dM, nacc);
      }
      __syncthreads();
    }
  }
  dS1 = wave_sum(dS1);
  dS2 = wave_sum(dS2);
  dM = wave_sum(dM);
  nacc = wave_sum(nacc);
  const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
  if (lane == 0) {
    red[0][wid] = dS1;
    red[1][wid] = dS2;
    red[2][wid] = dM;
    red[3][wid] = nacc;
  }
  __syncthreads();
  if (t == 0) {
    long long a1 = 0, a2 = 0, am = 0, an = 0;
    for (int w = 0; w < PAIR_BLOCK / 64; ++w) {
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

// LDS-tiled line-class kernel (16-site chunk path).  Work items are (chain,
// jh0, kh0, TJ | TK << 16): the TJ x TK class lines j = cj + 2*jh,
// k = ck + 2*kh with jh in [jh0, jh0+TJ), kh in [kh0, kh0+TK) (2D: TK = 1,
// k = 0).  The (2TJ+1) x (2TK+1) lines around them -- every line their sites
// read -- are staged in LDS once, both halves are updated there and only the
// class lines are written back, so each lattice line is read about once per
// launch instead of once per class line that neighbours it.  Neighbour lines
// have the other j or k parity, so they are never class lines of this launch
// and tiles do not interact.  Workgroups are persistent: while a workgroup
// updates one tile, the lines of its next tile are already on their way into
// registers (TILE_PF = 12 16-byte words per thread, at most 48 KB of lines
// per tile).  Random counters are those of the per-colour kernels (same line
// index ln), so trajectories are identical.
constexpr int TILE_PF = 12;

struct TileGeom {
  int c, jh0, kh0, TJ, TK, RJ, RK, n16;
};

__device__ __forceinline__ TileGeom tile_geom(const CheckerParams *__restrict__ Pp,
                                              const int32_t *__restrict__ work, int it) {
  TileGeom t;
  const int32_t *item = work + 4 * it;
  t.c = item[0];
  t.jh0 = item[1];
  t.kh0 = item[2];
  t.TJ = item[3] & 0xFFFF;
  t.TK = item[3] >> 16;
  const bool d3 = Pp[t.c].ndim == 3;
  t.RJ = 2 * t.TJ + 1;
  t.RK = d3 ? 2 * t.TK + 1 : 1;
  t.n16 = t.RJ * t.RK * (Pp[t.c].L0 >> 4);
  return t;
}

__global__ __launch_bounds__(PAIR_BLOCK) void checker_tile_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ work, int n_items, int line_class) {
  extern __shared__ __align__(16) uint8_t region[];
  __shared__ int red[4][PAIR_BLOCK / 64];
  __shared__ uint32_t tab[CHECKER_NTAB];  // acceptance table of the current tile's chain
  uint4 *reg16 = reinterpret_cast<uint4 *>(region);
  const int cj = line_class & 1, ck = (line_class >> 1) & 1;
  // prefetch words as named registers (an array of this size would live in scratch)
  uint4 pf0, pf1, pf2, pf3, pf4, pf5, pf6, pf7, pf8, pf9, pf10, pf11;
// lines of tile ITEM into pf (thread t holds words t + r*PAIR_BLOCK)
#define MCG_TILE_FETCH(ITEM)                                                                   \
  do {                                                                                         \
    const TileGeom ft = tile_geom(Pp, work, (ITEM));                                           \
    const CheckerParams &Q = Pp[ft.c];                                                         \
    const uint8_t *qlat = reinterpret_cast<const uint8_t *>(lat_pool + Q.lat_off);             \
    const int qv16 = Q.L0 >> 4;                                                                \
    const bool q3 = Q.ndim == 3;                                                               \
    const int qjb = cj + 2 * ft.jh0 - 1, qkb = q3 ? ck + 2 * ft.kh0 - 1 : 0;                   \
    const int64_t qplane = (int64_t)Q.L0 * Q.L1;                                               \
    auto ld = [&](int r) -> uint4 {                                                            \
      const int idx = threadIdx.x + r * PAIR_BLOCK;                                            \
      if (idx >= ft.n16) return make_uint4(0, 0, 0, 0);                                        \
      const int row = idx / qv16, col = idx - row * qv16;                                      \
      const int dk = row / ft.RJ, dj = row - dk * ft.RJ;                                       \
      const int jj = wrapm(qjb + dj, Q.L1), kk = q3 ? wrapm(qkb + dk, Q.L2) : 0;               \
      return *reinterpret_cast<const uint4 *>(qlat + kk * qplane + (int64_t)jj * Q.L0 +       \
                                              col * 16);                                       \
    };                                                                                         \
    pf0 = ld(0); pf1 = ld(1); pf2 = ld(2); pf3 = ld(3); pf4 = ld(4); pf5 = ld(5);             \
    pf6 = ld(6); pf7 = ld(7); pf8 = ld(8); pf9 = ld(9); pf10 = ld(10); pf11 = ld(11);         \
  } while (0)
// pf into the LDS region of tile ITEM, and its chain's acceptance table
#define MCG_TILE_STAGE(ITEM)                                                                   \
  do {                                                                                         \
    const TileGeom sg = tile_geom(Pp, work, (ITEM));                                           \
    const int sn16 = sg.n16;                                                                   \
    for (int i = threadIdx.x; i < CHECKER_NTAB; i += PAIR_BLOCK)                               \
      tab[i] = tables[Pp[sg.c].tab_off + i];                                                   \
    auto st = [&](int r, const uint4 &v) {                                                     \
      const int idx = threadIdx.x + r * PAIR_BLOCK;                                            \
      if (idx < sn16) reg16[idx] = v;                                                          \
    };                                                                                         \
    st(0, pf0); st(1, pf1); st(2, pf2); st(3, pf3); st(4, pf4); st(5, pf5);                   \
    st(6, pf6); st(7, pf7); st(8, pf8); st(9, pf9); st(10, pf10); st(11, pf11);               \
  } while (0)

  int it = blockIdx.x;
  if (it < n_items) {
    MCG_TILE_FETCH(it);
    MCG_TILE_STAGE(it);
  }
  __syncthreads();
  while (it < n_items) {
    const int nxt = it + (int)gridDim.x;
    if (nxt < n_items) MCG_TILE_FETCH(nxt);  // in flight while this tile is updated
    const TileGeom t = tile_geom(Pp, work, it);
    const CheckerParams P = Pp[t.c];
    CheckerState *St = Sp + t.c;
    const LineGeom g = line_geom(P);
    const int64_t pass = St->n_pass;
    uint8_t *lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
    const int upl = g.half >> 4;
    const bool d3 = P.ndim == 3;
    const int L0 = P.L0, v16 = L0 >> 4, h1 = P.L1 >> 1;
    const int64_t plane = (int64_t)L0 * P.L1;
    auto line_at = [&](int dj, int dk) { return region + (dj + t.RJ * dk) * L0; };
    int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
    for (int a = 0; a < 2; ++a) {
      const int color = 2 * line_class + a;
      for (int task = threadIdx.x; task < t.TJ * t.TK * upl; task += PAIR_BLOCK) {
        const int li = task / upl, u = task - li * upl;
        const int tk = li / t.TJ, tj = li - tk * t.TJ;
        const int ln = (t.jh0 + tj) + h1 * (t.kh0 + tk);
        const int dj = 2 * tj + 1, dk = d3 ? 2 * tk + 1 : 0;
        Lines Ls;
        Ls.own = line_at(dj, dk);
        Ls.jp = line_at(dj + 1, dk);
        Ls.jm = line_at(dj - 1, dk);
        Ls.kp = Ls.km = Ls.jpkp = Ls.jpkm = Ls.jmkp = Ls.jmkm = Ls.own;
        if (d3) {
          Ls.kp = line_at(dj, dk + 1);
          Ls.km = line_at(dj, dk - 1);
          Ls.jpkp = line_at(dj + 1, dk + 1);
          Ls.jpkm = line_at(dj + 1, dk - 1);
          Ls.jmkp = line_at(dj - 1, dk + 1);
          Ls.jmkm = line_at(dj - 1, dk - 1);
        }
        update_chunk_at(P, g, Ls, a, ln, u << 4, tab, color, pass, dS1, dS2, dM, nacc);
      }
      __syncthreads();
    }
    for (int idx = threadIdx.x; idx < t.TJ * t.TK * v16; idx += PAIR_BLOCK) {
      const int li = idx / v16, col = idx - li * v16;
      const int tk = li / t.TJ, tj = li - tk * t.TJ;
      const int j = cj + 2 * (t.jh0 + tj), k = d3 ? ck + 2 * (t.kh0 + tk) : 0;
      *reinterpret_cast<uint4 *>(lat + k * plane + (int64_t)j * L0 + col * 16) =
          reinterpret_cast<const uint4 *>(line_at(2 * tj + 1, d3 ? 2 * tk + 1 : 0))[col];
    }
    dS1 = wave_sum(dS1);
    dS2 = wave_sum(dS2);
    dM = wave_sum(dM);
    nacc = wave_sum(nacc);
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    if (lane == 0) {
      red[0][wid] = dS1;
      red[1][wid] = dS2;
      red[2][wid] = dM;
      red[3][wid] = nacc;
    }
    __syncthreads();  // also: every read of this tile's lines is done
    if (tid == 0) {
      long long a1 = 0, a2 = 0, am = 0, an = 0;
      for (int w = 0; w < PAIR_BLOCK / 64; ++w) {
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
    if (nxt < n_items) MCG_TILE_STAGE(nxt);
    __syncthreads();
    it = nxt;
  }
#undef MCG_TILE_FETCH
#undef MCG_TILE_STAGE
}

// ---------------------------------------------------------------------------
// 2D band kernels (HBM path, 16-site chunk layout): one pass in two launches
// that read and write every line about once, instead of one launch per colour
// that each read the whole lattice.
//
// A pass splits into two waves of updates: A (2 colours: colour 0 on every
// line; 4 colours: colours 0 then 1 on the even lines) and B (colour 1 on
// every line; colours 2 then 3 on the odd lines).  A(j) reads the B sites of
// lines j-1..j+1 before the pass touches them; B(j) needs A done on
// j-1..j+1.  A workgroup owns a band of H lines [j0, j0+H) and streams it
// through an LDS ring of BAND_RING lines: step t runs A(t+1) beside
// B(t-1) (2 colours) or B(t-2) (4 colours, odd t only), pairs that neither
// read nor write each other's sites.  B on the band's edge lines j0 and
// j0+H-1 needs lines of the neighbouring bands and runs in the second launch
// (checker_seam_kernel), from HBM.  Every site's update reads the same
// neighbour values as in the colour-by-colour sweep and uses the same Philox
// words, so trajectories are identical to the per-colour kernels.
//
// Write-back: a line goes back to HBM once its last update of the band is
// done, and only the halves the band changed (the edge lines hold only A
// updates), so no band writes bytes another band reads.  A band's ring copies
// of lines j0-1 and j0+H are read-only; of them only the halves holding B
// sites are used (the other halves may be changing in the neighbouring band).
// ---------------------------------------------------------------------------
constexpr int BAND_BLOCK = 256;
constexpr int BAND_RING = 8;

struct BandCtx {
  const CheckerParams *P;
  LineGeom g;
  uint8_t *ring;
  int L0, L1, j0, H;
  __device__ __forceinline__ uint8_t *slot(int j) const {
    return ring + ((j - j0 + 1) & (BAND_RING - 1)) * L0;
  }
};

// One colour's half-line update of line j (all its chunks), by threads
// [t0, t0 + nthreads) of the block.
__device__ __forceinline__ void band_half(const BandCtx &b, int j, int color, int t0,
                                          int nthreads, const uint32_t *tab, int64_t pass,
                                          int &dS1, int &dS2, int &dM, int &nacc) {
  const int cpl = b.g.half >> 4;
  const int a = b.P->ncolor == 2 ? ((color + j) & 1) : (color & 1);
  const int ln = b.P->ncolor == 2 ? j : (j >> 1);
  Lines Ls;
  Ls.own = b.slot(j);
  Ls.jp = b.slot(j + 1);
  Ls.jm = b.slot(j - 1);
  Ls.kp = Ls.km = Ls.jpkp = Ls.jpkm = Ls.jmkp = Ls.jmkm = Ls.own;
  for (int ch = (int)threadIdx.x - t0; ch < cpl; ch += nthreads)
    update_chunk_at(*b.P, b.g, Ls, a, ln, ch << 4, tab, color, pass, dS1, dS2, dM, nacc);
}

// Lines of the band: copy line j (global, wrapped) into its ring slot.
__device__ __forceinline__ void band_load(const BandCtx &b, const uint8_t *lat, int j) {
  const int jw = j < 0 ? j + b.L1 : (j >= b.L1 ? j - b.L1 : j);
  const uint4 *src = reinterpret_cast<const uint4 *>(lat + (int64_t)jw * b.L0);
  uint4 *dst = reinterpret_cast<uint4 *>(b.slot(j));
  for (int i = threadIdx.x; i < (b.L0 >> 4); i += BAND_BLOCK) dst[i] = src[i];
}

// Write back half `which` (0, 1) or both halves (2) of band line j.
__device__ __forceinline__ void band_store(const BandCtx &b, uint8_t *lat, int j, int which) {
  const int n16 = b.g.half >> 4;
  const int first = which == 1 ? n16 : 0;
  const int count = which == 2 ? 2 * n16 : n16;
  const uint4 *src = reinterpret_cast<const uint4 *>(b.slot(j));
  uint4 *dst = reinterpret_cast<uint4 *>(lat + (int64_t)j * b.L0);
  for (int i = threadIdx.x; i < count; i += BAND_BLOCK) dst[first + i] = src[first + i];
}

__device__ __forceinline__ void block_add_deltas(CheckerState *St, int dS1, int dS2, int dM,
                                                 int nacc, int (*red)[BAND_BLOCK / 64]) {
  dS1 = wave_sum(dS1);
  dS2 = wave_sum(dS2);
  dM = wave_sum(dM);
  nacc = wave_sum(nacc);
  const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
  if (lane == 0) {
    red[0][wid] = dS1;
    red[1][wid] = dS2;
    red[2][wid] = dM;
    red[3][wid] = nacc;
  }
  __syncthreads();
  if (t == 0) {
    long long a1 = 0, a2 = 0, am = 0, an = 0;
    for (int w = 0; w < BAND_BLOCK / 64; ++w) {
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

// work: (chain, j0, H) per band
__global__ __launch_bounds__(BAND_BLOCK) void checker_band_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ work) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t tab[CHECKER_NTAB];
  __shared__ int red[4][BAND_BLOCK / 64];
  const int c = work[3 * blockIdx.x], j0 = work[3 * blockIdx.x + 1], H = work[3 * blockIdx.x + 2];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  const int64_t pass = St->n_pass;
  for (int i = threadIdx.x; i < CHECKER_NTAB; i += BAND_BLOCK) tab[i] = tables[P.tab_off + i];
  uint8_t *lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  BandCtx b;
  b.P = &P;
  b.g = line_geom(P);
  b.ring = reinterpret_cast<uint8_t *>(smem);
  b.L0 = P.L0;
  b.L1 = P.L1;
  b.j0 = j0;
  b.H = H;
  const bool four = P.ncolor == 4;
  const int stride = four ? 2 : 1;
  for (int j = j0 - 1; j <= j0 + 1; ++j) band_load(b, lat, j);
  __syncthreads();
  int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
  const int jlast = j0 + H - 1;
  for (int t = j0 - 1; t <= jlast; t += stride) {
    if (!four) {
      const bool doA = t + 1 <= jlast;
      const bool doB = t - 1 >= j0 + 1 && t - 1 <= jlast - 1;
      // A(t+1) on threads [0, nA), B(t-1) on the rest
      const int nA = doA ? (doB ? BAND_BLOCK / 2 : BAND_BLOCK) : 0;
      if (doA && (int)threadIdx.x < nA)
        band_half(b, t + 1, 0, 0, nA, tab, pass, dS1, dS2, dM, nacc);
      if (doB && (int)threadIdx.x >= nA)
        band_half(b, t - 1, 1, nA, BAND_BLOCK - nA, tab, pass, dS1, dS2, dM, nacc);
    } else {
      const bool doA = t + 1 <= jlast - 1;
      const bool doB = t - 2 >= j0 + 1 && t - 2 <= jlast - 2;
      const int nA = doA ? (doB ? BAND_BLOCK / 2 : BAND_BLOCK) : 0;
      if (doA && (int)threadIdx.x < nA)
        band_half(b, t + 1, 0, 0, nA, tab, pass, dS1, dS2, dM, nacc);
      if (doB && (int)threadIdx.x >= nA)
        band_half(b, t - 2, 2, nA, BAND_BLOCK - nA, tab, pass, dS1, dS2, dM, nacc);
      __syncthreads();
      if (doA && (int)threadIdx.x < nA)
        band_half(b, t + 1, 1, 0, nA, tab, pass, dS1, dS2, dM, nacc);
      if (doB && (int)threadIdx.x >= nA)
        band_half(b, t - 2, 3, nA, BAND_BLOCK - nA, tab, pass, dS1, dS2, dM, nacc);
    }
    // the next lines into their ring slots (slots of lines no step reads any more)
    for (int j = t + 3; j < t + 3 + stride; ++j)
      if (j <= jlast + 1) band_load(b, lat, j);
    __syncthreads();
    // lines whose last update of this band was this step
    if (!four) {
      // A-only edge lines j0 and jlast: the A half; inner lines: after B(j) at step j+1
      if (t + 1 == j0 || (t + 1 == jlast && jlast != j0)) band_store(b, lat, t + 1, (t + 1) & 1);
      if (t - 1 >= j0 + 1 && t - 1 <= jlast - 1) band_store(b, lat, t - 1, 2);
    } else {
      if (t + 1 <= jlast - 1) band_store(b, lat, t + 1, 2);           // even line after A
      if (t - 2 >= j0 + 1 && t - 2 <= jlast - 2) band_store(b, lat, t - 2, 2);  // odd after B
    }
  }
  block_add_deltas(St, dS1, dS2, dM, nacc, red);
}

// B on the band edge lines (2 colours: j0 and j0+H-1; 4 colours: j0+H-1),
// from HBM after every band's first launch.
__global__ __launch_bounds__(BAND_BLOCK) void checker_seam_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ work) {
  __shared__ int red[4][BAND_BLOCK / 64];
  const int c = work[3 * blockIdx.x], j0 = work[3 * blockIdx.x + 1], H = work[3 * blockIdx.x + 2];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  const int64_t pass = St->n_pass;
  const uint32_t *tab = tables + P.tab_off;
  uint8_t *lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  const LineGeom g = line_geom(P);
  const int cpl = g.half >> 4;
  int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
  const int jlast = j0 + H - 1;
  if (P.ncolor == 2) {
    const int nl = jlast != j0 ? 2 : 1;
    for (int task = threadIdx.x; task < nl * cpl; task += BAND_BLOCK) {
      const int j = task < cpl ? j0 : jlast;
      const int ch = task < cpl ? task : task - cpl;
      update_chunk(P, g, lat, tab, 1, pass, (int64_t)j * cpl + ch, dS1, dS2, dM, nacc);
    }
  } else {
    for (int ch = threadIdx.x; ch < cpl; ch += BAND_BLOCK)
      update_chunk(P, g, lat, tab, 2, pass, (int64_t)(jlast >> 1) * cpl + ch, dS1, dS2, dM, nacc);
    __syncthreads();
    for (int ch = threadIdx.x; ch < cpl; ch += BAND_BLOCK)
      update_chunk(P, g, lat, tab, 3, pass, (int64_t)(jlast >> 1) * cpl + ch, dS1, dS2, dM, nacc);
  }
  block_add_deltas(St, dS1, dS2, dM, nacc, red);
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
      mcg_sample rec = {};
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
                          int color, int upt, hipStream_t stream) {
  if (upt == 1)
    hipLaunchKernelGGL(checker_color_kernel<1>, dim3(n_work_items * 4), dim3(HBM_BLOCK), 0,
                       stream, P, St, lat_pool, tables, work, color);
  else if (upt == 2)
    hipLaunchKernelGGL(checker_color_kernel<2>, dim3(n_work_items * 2), dim3(HBM_BLOCK), 0,
                       stream, P, St, lat_pool, tables, work, color);
  else
    hipLaunchKernelGGL(checker_color_kernel<4>, dim3(n_work_items), dim3(HBM_BLOCK), 0, stream,
                       P, St, lat_pool, tables, work, color);
}

void launch_checker_record(const CheckerParams *P, CheckerState *St, mcg_sample *series,
                           int64_t series_cap, const uint8_t *active, int n_chains,
                           int64_t sample_period, hipStream_t stream) {
  hipLaunchKernelGGL(checker_record_kernel, dim3((n_chains + 63) / 64), dim3(64), 0, stream, P,
                     St, series, series_cap, active, n_chains, sample_period);
}

// `passes` passes of the HBM path as a chain of kernel nodes of graph g (an
// explicitly built graph: no stream capture, so building one never interacts
// with work other threads issue meanwhile).
hipError_t checker_graph_add_passes(hipGraph_t g, const CheckerParams *P, CheckerState *St,
                                    int8_t *lat_pool, const uint32_t *tables,
                                    const int32_t *work, int n_work_items, int ncolor, int upt,
                                    mcg_sample *series, int64_t series_cap,
                                    const uint8_t *active, int n_chains, int64_t sample_period,
                                    int passes, const int32_t *band_work, int n_bands,
                                    int ring_bytes, bool slabs, const int32_t *pair_work,
                                    int n_pairs, int pair_lds, int tile_grid) {
  void *color_fn = upt == 1   ? reinterpret_cast<void *>(&checker_color_kernel<1>)
                   : upt == 2 ? reinterpret_cast<void *>(&checker_color_kernel<2>)
                              : reinterpret_cast<void *>(&checker_color_kernel<4>);
  const unsigned color_grid = (unsigned)n_work_items * (unsigned)(4 / upt);
  hipGraphNode_t prev = nullptr;
  auto add = [&](void *fn, dim3 grid, dim3 block, void **args, unsigned shmem = 0) -> hipError_t {
    hipKernelNodeParams kp{};
    kp.func = fn;
    kp.gridDim = grid;
    kp.blockDim = block;
    kp.sharedMemBytes = shmem;
    kp.kernelParams = args;
    kp.extra = nullptr;
    hipGraphNode_t node = nullptr;
    hipError_t e = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
    prev = node;
    return e;
  };
  for (int p = 0; p < passes; ++p) {
    if (n_bands > 0 && slabs) {
      void *args[] = {(void *)&P, (void *)&St, (void *)&lat_pool, (void *)&tables,
                      (void *)&band_work};
      const unsigned blk = (unsigned)checker_slab_block();
      hipError_t e = add(checker_slab_kernel_fn(), dim3(n_bands), dim3(blk), args,
                         (unsigned)ring_bytes);
      if (e == hipSuccess)
        e = add(checker_slab_seam_kernel_fn(), dim3(n_bands), dim3(blk), args,
                (unsigned)(3 * ring_bytes / checker_slab_ring_planes()));
      if (e != hipSuccess) return e;
    } else if (n_bands > 0) {
      void *args[] = {(void *)&P, (void *)&St, (void *)&lat_pool, (void *)&tables,
                      (void *)&band_work};
      hipError_t e = add(reinterpret_cast<void *>(&checker_band_kernel), dim3(n_bands),
                         dim3(BAND_BLOCK), args, (unsigned)ring_bytes);
      if (e == hipSuccess)
        e = add(reinterpret_cast<void *>(&checker_seam_kernel), dim3(n_bands), dim3(BAND_BLOCK),
                args);
      if (e != hipSuccess) return e;
    } else if (n_pairs > 0) {
      for (int lc = 0; lc < ncolor / 2; ++lc) {
        hipError_t e;
        if (pair_lds > 0) {
          void *args[] = {(void *)&P,         (void *)&St,      (void *)&lat_pool,
                          (void *)&tables,    (void *)&pair_work, (void *)&n_pairs,
                          (void *)&lc};
          e = add(reinterpret_cast<void *>(&checker_tile_kernel), dim3(tile_grid),
                  dim3(PAIR_BLOCK), args, (unsigned)pair_lds);
        } else {
          void *args[] = {(void *)&P, (void *)&St, (void *)&lat_pool, (void *)&tables,
                          (void *)&pair_work, (void *)&lc};
          e = add(reinterpret_cast<void *>(&checker_pair_kernel), dim3(n_pairs),
                  dim3(PAIR_BLOCK), args);
        }
        if (e != hipSuccess) return e;
      }
    } else if (n_work_items > 0)
      for (int color = 0; color < ncolor; ++color) {
        void *args[] = {(void *)&P, (void *)&St, (void *)&lat_pool, (void *)&tables,
                        (void *)&work, (void *)&color};
        hipError_t e = add(color_fn, dim3(color_grid), dim3(HBM_BLOCK), args);
        if (e != hipSuccess) return e;
      }
    void *args[] = {(void *)&P,          (void *)&St,     (void *)&series,
                    (void *)&series_cap, (void *)&active, (void *)&n_chains,
                    (void *)&sample_period};
    hipError_t e = add(reinterpret_cast<void *>(&checker_record_kernel),
                       dim3((n_chains + 63) / 64), dim3(64), args);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void launch_checker_band(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                         const uint32_t *tables, const int32_t *band_work, int n_bands,
                         int ring_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(checker_band_kernel, dim3(n_bands), dim3(BAND_BLOCK), ring_bytes, stream,
                     P, St, lat_pool, tables, band_work);
  hipLaunchKernelGGL(checker_seam_kernel, dim3(n_bands), dim3(BAND_BLOCK), 0, stream, P, St,
                     lat_pool, tables, band_work);
}

void launch_checker_pair(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                         const uint32_t *tables, const int32_t *pair_work, int n_items,
                         int line_class, int lds_bytes, int tile_grid, hipStream_t stream) {
  if (lds_bytes > 0)
    hipLaunchKernelGGL(checker_tile_kernel, dim3(tile_grid), dim3(PAIR_BLOCK), lds_bytes, stream,
                       P, St, lat_pool, tables, pair_work, n_items, line_class);
  else
    hipLaunchKernelGGL(checker_pair_kernel, dim3(n_items), dim3(PAIR_BLOCK), 0, stream, P, St,
                       lat_pool, tables, pair_work, line_class);
}

hipError_t checker_pair_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(checker_tile_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

int checker_tile_max_bytes() { return TILE_PF * PAIR_BLOCK * 16; }

int checker_pair_block() { return PAIR_BLOCK; }

int checker_band_ring_lines() { return BAND_RING; }

hipError_t checker_band_set_max_lds(int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(checker_band_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

int checker_units_per_item() { return UNITS_PER_ITEM; }

}  // namespace mcg
