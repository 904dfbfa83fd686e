// checker_slab.hip -- CHECKERBOARD HBM path of the 3D simple-cubic J1+J2 model
// (8 colours (i&1) + 2(j&1) + 4(k&1), split-row layout, half lines a multiple
// of 32 sites): a slab of planes per workgroup, the planes bit-packed in LDS.
//
// Order of updates.  A plane holds one k parity, so an even plane carries
// colours 0-3 and an odd plane colours 4-7; a site's neighbours lie in its own
// plane and the planes k+-1.  A pass splits into A (colours 0-3 on the even
// planes, in colour order) and B (colours 4-7 on the odd planes): A(k) reads
// the B sites of planes k+-1 before the pass touches them, B(k) needs A done
// on k+-1.  This is the 2D four-colour band scheme with planes for lines: a
// workgroup owns the planes [k0, k0+H) (k0 and H even) and streams them
// through an LDS ring of SLAB_RING planes; step t runs A(t+1) beside B(t-2),
// which neither read nor write each other's planes.  B on the slab's last
// plane needs the next slab's first plane after its A, so it runs in a
// second launch (checker_slab_seam_kernel).  Within a plane the four colours
// run in order with a barrier between them; a workgroup of 512 threads (two
// words per thread and colour on 256^2 planes) leaves room for two slabs per
// CU, whose barriers do not wait for each other.  Every site reads the neighbour
// values of the colour-by-colour sweep and the same Philox words (counters of
// colour-class line ln = (j>>1) + (L1/2)(k>>1), as in checker.hip), so the
// trajectory equals the per-colour and line-class kernels' bit for bit.
//
// Bit planes.  In LDS a row (line j of a plane) is its even-i half then its
// odd-i half, one bit per site, L0/64 words per half, rows padded to an odd
// word count.  A thread updates one 32-bit word of a colour: 32 sites of one
// half row, all of that colour.  Their 6 nearest and 12 next-nearest
// neighbour bits come as 18 words (the in-line and in-plane-diagonal ones
// shifted by one site with the carry bit from the next word), are added as
// bit vectors (carry-save adders: the counts c1 <= 6, c2 <= 12 as 3 and 4 bit
// planes), and the class byte (b << 7) | (c1 << 4) | c2 of four sites at a time
// comes from spreading nibbles of the bit planes into bytes.  The flip mask is
// xor-ed into the word; the integer observables change by popcounts of the
// flip mask against the count planes.  HBM keeps the byte lattice: planes are
// packed on load and unpacked on store, each plane read once per launch (plus
// the plane before the slab) and written once per pass.
#include "internal.hh"

namespace mcg {

namespace {

constexpr int SLAB_BLOCK = 512;
constexpr int SLAB_RING = 8;

struct SlabGeom {
  int L0, L1, L2;
  int half;  // sites per half line
  int hw;    // 32-bit words per half line
  int rs;    // LDS words per row (2 hw + 1)
  int pw;    // LDS words per plane
  int h1;    // L1 / 2
  int opl;   // Philox calls per half line (half / 8)
};

__device__ __forceinline__ SlabGeom slab_geom(const CheckerParams &P) {
  SlabGeom G;
  G.L0 = P.L0;
  G.L1 = P.L1;
  G.L2 = P.L2;
  G.half = P.L0 >> 1;
  G.hw = P.L0 >> 6;
  G.rs = 2 * G.hw + 1;
  G.pw = G.rs * P.L1;
  G.h1 = P.L1 >> 1;
  G.opl = G.half >> 3;
  return G;
}

// 4 bytes holding 0/1 -> 4 bits (byte e -> bit e): each byte's bit lands on
// its own position in 28..31 and no two partial products meet.
__device__ __forceinline__ uint32_t pack4(uint32_t v) { return (v * 0x10204080u) >> 28; }
// 4 bits -> 4 bytes holding 0/1 (bit e -> byte e)
__device__ __forceinline__ uint32_t unpack4(uint32_t x) {
  return (x * 0x00204081u) & 0x01010101u;
}

__device__ __forceinline__ uint32_t pack16(uint4 q) {
  return pack4(q.x) | (pack4(q.y) << 4) | (pack4(q.z) << 8) | (pack4(q.w) << 12);
}

// global byte address of the 32 sites of word w of half a of line (j, k)
__device__ __forceinline__ const uint8_t *site_bytes(const uint8_t *lat, const SlabGeom &G,
                                                     int j, int k, int a, int w) {
  return lat + ((int64_t)k * G.L1 + j) * G.L0 + a * G.half + 32 * w;
}

// Plane k (global, wrapped) -> LDS slot, by threads [t0, t0 + nt).
__device__ __forceinline__ void slab_load(const SlabGeom &G, const uint8_t *lat, int k,
                                          uint32_t *slot, int t0, int nt) {
  const int kw = k < 0 ? k + G.L2 : (k >= G.L2 ? k - G.L2 : k);
  const int words = G.L1 * 2 * G.hw;
  for (int wi = (int)threadIdx.x - t0; wi < words; wi += nt) {
    const int j = wi / (2 * G.hw), r = wi - j * 2 * G.hw;
    const int a = r >= G.hw ? 1 : 0, w = r - a * G.hw;
    const uint4 *src = reinterpret_cast<const uint4 *>(site_bytes(lat, G, j, kw, a, w));
    const uint4 q0 = src[0], q1 = src[1];
    slot[j * G.rs + a * G.hw + w] = pack16(q0) | (pack16(q1) << 16);
  }
}

// LDS slot -> plane k (global), by threads [t0, t0 + nt).
__device__ __forceinline__ void slab_store(const SlabGeom &G, uint8_t *lat, int k,
                                           const uint32_t *slot, int t0, int nt) {
  const int words = G.L1 * 2 * G.hw;
  for (int wi = (int)threadIdx.x - t0; wi < words; wi += nt) {
    const int j = wi / (2 * G.hw), r = wi - j * 2 * G.hw;
    const int a = r >= G.hw ? 1 : 0, w = r - a * G.hw;
    const uint32_t x = slot[j * G.rs + a * G.hw + w];
    uint4 *dst = reinterpret_cast<uint4 *>(const_cast<uint8_t *>(site_bytes(lat, G, j, k, a, w)));
    dst[0] = make_uint4(unpack4(x & 15u), unpack4((x >> 4) & 15u), unpack4((x >> 8) & 15u),
                        unpack4((x >> 12) & 15u));
    dst[1] = make_uint4(unpack4((x >> 16) & 15u), unpack4((x >> 20) & 15u),
                        unpack4((x >> 24) & 15u), unpack4(x >> 28));
  }
}

// carry-save adder on bit vectors: s = x ^ y ^ z, c = majority(x, y, z)
__device__ __forceinline__ void csa(uint32_t x, uint32_t y, uint32_t z, uint32_t &s,
                                    uint32_t &c) {
  const uint32_t u = x ^ y;
  s = u ^ z;
  c = (x & y) | (u & z);
}

// nibble n of x spread to the low bit of four bytes
__device__ __forceinline__ uint32_t spread(uint32_t x, int n) {
  return __umul24((x >> (4 * n)) & 15u, 0x00204081u) & 0x01010101u;
}

struct SlabDelta {
  int dS1 = 0, dS2 = 0, dM = 0, nacc = 0;
};

// One word of colour `color` (half a of row j) of the plane in slot `own`,
// planes k-1 and k+1 in slots km and kp; k is the plane's index (Philox
// counters), pass the pass being made.
__device__ __forceinline__ void slab_word(const SlabGeom &G, const CheckerParams &P,
                                          uint32_t *own, const uint32_t *km, const uint32_t *kp,
                                          int j, int k, int a, int w, int color,
                                          const uint32_t *tab, uint32_t pass_lo,
                                          uint32_t pass_hi, SlabDelta &d) {
  const int o = a ^ 1;
  const int jp = j + 1 == G.L1 ? 0 : j + 1, jm = j == 0 ? G.L1 - 1 : j - 1;
  // the word that supplies the bit shifted in: a = 0 takes index h-1 (word
  // w-1, its top bit), a = 1 index h+1 (word w+1, its bottom bit)
  const int wn = a ? (w + 1 == G.hw ? 0 : w + 1) : (w == 0 ? G.hw - 1 : w - 1);
  const int rj = j * G.rs, rjp = jp * G.rs, rjm = jm * G.rs;
  const int sa = a * G.hw + w, so = o * G.hw + w, sn = o * G.hw + wn;
  auto sh = [&](uint32_t x, uint32_t xn) -> uint32_t {
    return a ? ((x >> 1) | (xn << 31)) : ((x << 1) | (xn >> 31));
  };
  const uint32_t b = own[rj + sa];
  // nearest neighbours: in-line (other half, indices h and h-1+2a), rows j+-1,
  // planes k+-1
  const uint32_t oo = own[rj + so], oon = own[rj + sn];
  const uint32_t n1 = oo, n2 = sh(oo, oon);
  const uint32_t n3 = own[rjp + sa], n4 = own[rjm + sa];
  const uint32_t n5 = kp[rj + sa], n6 = km[rj + sa];
  // next-nearest: in-plane diagonals (rows j+-1, other half), (i+-1, k+-1)
  // (planes k+-1, row j, other half), (j+-1, k+-1) (planes k+-1, rows j+-1)
  const uint32_t pj = own[rjp + so], pjn = own[rjp + sn];
  const uint32_t mj = own[rjm + so], mjn = own[rjm + sn];
  const uint32_t pk = kp[rj + so], pkn = kp[rj + sn];
  const uint32_t mk = km[rj + so], mkn = km[rj + sn];
  const uint32_t m1 = pj, m2 = sh(pj, pjn), m3 = mj, m4 = sh(mj, mjn);
  const uint32_t m5 = pk, m6 = sh(pk, pkn), m7 = mk, m8 = sh(mk, mkn);
  const uint32_t m9 = kp[rjp + sa], m10 = kp[rjm + sa], m11 = km[rjp + sa], m12 = km[rjm + sa];
  // c1 = n1 + ... + n6 as bit planes c1_0..c1_2
  uint32_t sa1, ca1, sb1, cb1, c1_0, c1_1, c1_2;
  csa(n1, n2, n3, sa1, ca1);
  csa(n4, n5, n6, sb1, cb1);
  c1_0 = sa1 ^ sb1;
  csa(ca1, cb1, sa1 & sb1, c1_1, c1_2);
  // c2 = m1 + ... + m12 as bit planes c2_0..c2_3
  uint32_t s1, k1, s2, k2, s3, k3, s4, k4, u1, v1;
  csa(m1, m2, m3, s1, k1);
  csa(m4, m5, m6, s2, k2);
  csa(m7, m8, m9, s3, k3);
  csa(m10, m11, m12, s4, k4);
  csa(s1, s2, s3, u1, v1);
  const uint32_t c2_0 = u1 ^ s4, w1 = u1 & s4;
  uint32_t x1, y1, x2, y2, c2_2, c2_3;
  csa(k1, k2, k3, x1, y1);
  csa(k4, v1, w1, x2, y2);
  const uint32_t c2_1 = x1 ^ x2;
  csa(y1, y2, x1 & x2, c2_2, c2_3);

  // random words: class line ln, Philox call ln*opl + h/8 per 8 sites
  const int ln = (j >> 1) + G.h1 * (k >> 1);
  const uint32_t ctr0 = (uint32_t)((int64_t)ln * G.opl + 4 * w);
  uint32_t flip = 0u, tie = 0u;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const u32x4 r = philox4x32_10(ctr0 + q, pass_lo, pass_hi, (uint32_t)color, P.key0, P.key1);
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      const int n = 2 * q + nb;  // nibble: sites 4n .. 4n+3
      const uint32_t cls = (spread(b, n) << 7) | (spread(c1_2, n) << 6) |
                           (spread(c1_1, n) << 5) | (spread(c1_0, n) << 4) |
                           (spread(c2_3, n) << 3) | (spread(c2_2, n) << 2) |
                           (spread(c2_1, n) << 1) | spread(c2_0, n);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int p = 4 * n + e;           // site h = 32w + p; within its 8-site group e8
        const int e8 = 4 * nb + e;
        const uint32_t word = r.v[e8 >> 1];
        const uint32_t hi16 = (e8 & 1) ? (word >> 16) : (word & 0xFFFFu);
        const uint32_t t = tab[(cls >> (8 * e)) & 0xFFu];
        const uint32_t thi = t >> 16;
        flip |= (uint32_t)(t == 0xFFFFFFFFu || hi16 < thi) << p;
        tie |= (uint32_t)(t != 0xFFFFFFFFu && hi16 == thi) << p;
      }
    }
  }
  while (tie) {  // rare: the high halves tie, the refinement word decides
    const int p = __builtin_ctz(tie);
    tie &= tie - 1u;
    const uint32_t cls = (((b >> p) & 1u) << 7) | (((c1_2 >> p) & 1u) << 6) |
                         (((c1_1 >> p) & 1u) << 5) | (((c1_0 >> p) & 1u) << 4) |
                         (((c2_3 >> p) & 1u) << 3) | (((c2_2 >> p) & 1u) << 2) |
                         (((c2_1 >> p) & 1u) << 1) | ((c2_0 >> p) & 1u);
    const uint32_t t = tab[cls];
    const uint32_t site_ctr = (uint32_t)((int64_t)ln * G.half + 32 * w + p);
    const uint32_t lo16 = philox4x32_10(site_ctr, pass_lo, pass_hi, (uint32_t)color | 0x80000000u,
                                        P.key0, P.key1)
                              .v[0] & 0xFFFFu;
    if (lo16 <= (t & 0xFFFFu)) flip |= 1u << p;
  }
  own[rj + sa] = b ^ flip;
  // deltas (count_flip summed): s = 2b - 1, dS1 = -2 s (2 c1 - z1),
  // dS2 = -2 s (2 c2 - z2), dM = -2 s
  const uint32_t f1 = flip & b, f0 = flip & ~b;
  const int n1c = __popc(f1), n0c = __popc(f0);
  const int sc1_1 = __popc(f1 & c1_0) + 2 * __popc(f1 & c1_1) + 4 * __popc(f1 & c1_2);
  const int sc1_0 = __popc(f0 & c1_0) + 2 * __popc(f0 & c1_1) + 4 * __popc(f0 & c1_2);
  const int sc2_1 = __popc(f1 & c2_0) + 2 * __popc(f1 & c2_1) + 4 * __popc(f1 & c2_2) +
                    8 * __popc(f1 & c2_3);
  const int sc2_0 = __popc(f0 & c2_0) + 2 * __popc(f0 & c2_1) + 4 * __popc(f0 & c2_2) +
                    8 * __popc(f0 & c2_3);
  d.dS1 += -4 * sc1_1 + 2 * P.z1 * n1c + 4 * sc1_0 - 2 * P.z1 * n0c;
  d.dS2 += -4 * sc2_1 + 2 * P.z2 * n1c + 4 * sc2_0 - 2 * P.z2 * n0c;
  d.dM += 2 * n0c - 2 * n1c;
  d.nacc += n0c + n1c;
}

// Colour `color` of plane k (slot own; planes k+-1 in km, kp) by threads
// [t0, t0 + nt).
__device__ __forceinline__ void slab_colour(const SlabGeom &G, const CheckerParams &P,
                                            uint32_t *own, const uint32_t *km,
                                            const uint32_t *kp, int k, int color,
                                            const uint32_t *tab, uint32_t pass_lo,
                                            uint32_t pass_hi, int t0, int nt, SlabDelta &d) {
  const int a = color & 1, rp = (color >> 1) & 1;
  const int words = G.h1 * G.hw;
  for (int wi = (int)threadIdx.x - t0; wi < words; wi += nt) {
    const int r = wi / G.hw, w = wi - r * G.hw;
    slab_word(G, P, own, km, kp, rp + 2 * r, k, a, w, color, tab, pass_lo, pass_hi, d);
  }
}

__device__ __forceinline__ void slab_add_deltas(CheckerState *St, SlabDelta d,
                                                long long (*red)[SLAB_BLOCK / 64]) {
  int v[4] = {d.dS1, d.dS2, d.dM, d.nacc};
#pragma unroll
  for (int q = 0; q < 4; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off);
  const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
  if (lane == 0)
    for (int q = 0; q < 4; ++q) red[q][wid] = v[q];
  __syncthreads();
  if (t == 0) {
    long long s[4] = {0, 0, 0, 0};
    for (int ww = 0; ww < SLAB_BLOCK / 64; ++ww)
      for (int q = 0; q < 4; ++q) s[q] += red[q][ww];
    if (s[0]) atomicAdd(reinterpret_cast<unsigned long long *>(&St->S1), (unsigned long long)s[0]);
    if (s[1]) atomicAdd(reinterpret_cast<unsigned long long *>(&St->S2), (unsigned long long)s[1]);
    if (s[2]) atomicAdd(reinterpret_cast<unsigned long long *>(&St->M), (unsigned long long)s[2]);
    if (s[3])
      atomicAdd(reinterpret_cast<unsigned long long *>(&St->n_accept), (unsigned long long)s[3]);
  }
}

// work: (chain, k0, H) per slab
__global__ __launch_bounds__(SLAB_BLOCK) void checker_slab_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ work) {
  extern __shared__ __align__(16) uint32_t ring[];
  __shared__ uint32_t tab[CHECKER_NTAB];
  __shared__ long long red[4][SLAB_BLOCK / 64];
  const int c = work[3 * blockIdx.x], k0 = work[3 * blockIdx.x + 1],
            H = work[3 * blockIdx.x + 2];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  const int64_t pass = St->n_pass;
  const uint32_t plo = (uint32_t)pass, phi = (uint32_t)(pass >> 32);
  for (int i = threadIdx.x; i < CHECKER_NTAB; i += SLAB_BLOCK) tab[i] = tables[P.tab_off + i];
  uint8_t *lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  const SlabGeom G = slab_geom(P);
  auto slot = [&](int k) { return ring + ((k - k0 + 1) & (SLAB_RING - 1)) * G.pw; };
  for (int k = k0 - 1; k <= k0 + 1; ++k) slab_load(G, lat, k, slot(k), 0, SLAB_BLOCK);
  __syncthreads();
  SlabDelta d;
  const int klast = k0 + H - 1;
  for (int t = k0 - 1; t <= klast; t += 2) {
    const bool doA = t + 1 <= klast - 1;
    const bool doB = t - 2 >= k0 + 1 && t - 2 <= klast - 2;
    const int nA = doA ? (doB ? SLAB_BLOCK / 2 : SLAB_BLOCK) : 0;
    const bool mineA = doA && (int)threadIdx.x < nA;
    const bool mineB = doB && (int)threadIdx.x >= nA;
    for (int cc = 0; cc < 4; ++cc) {
      if (mineA)
        slab_colour(G, P, slot(t + 1), slot(t), slot(t + 2), t + 1, cc, tab, plo, phi, 0, nA, d);
      if (mineB)
        slab_colour(G, P, slot(t - 2), slot(t - 3), slot(t - 1), t - 2, 4 + cc, tab, plo, phi,
                    nA, SLAB_BLOCK - nA, d);
      __syncthreads();
    }
    // the planes of the next step into the free slots, the finished planes out
    for (int k = t + 3; k <= t + 4 && k <= klast; ++k)
      slab_load(G, lat, k, slot(k), 0, SLAB_BLOCK);
    if (doA) slab_store(G, lat, t + 1, slot(t + 1), 0, SLAB_BLOCK);
    if (doB) slab_store(G, lat, t - 2, slot(t - 2), 0, SLAB_BLOCK);
    __syncthreads();
  }
  slab_add_deltas(St, d, red);
}

// B (colours 4-7) on the last plane of every slab, after every slab's first
// launch: planes klast-1, klast, klast+1 from HBM.
__global__ __launch_bounds__(SLAB_BLOCK) void checker_slab_seam_kernel(
    const CheckerParams *__restrict__ Pp, CheckerState *__restrict__ Sp,
    int8_t *__restrict__ lat_pool, const uint32_t *__restrict__ tables,
    const int32_t *__restrict__ work) {
  extern __shared__ __align__(16) uint32_t ring[];
  __shared__ uint32_t tab[CHECKER_NTAB];
  __shared__ long long red[4][SLAB_BLOCK / 64];
  const int c = work[3 * blockIdx.x], k0 = work[3 * blockIdx.x + 1],
            H = work[3 * blockIdx.x + 2];
  const CheckerParams P = Pp[c];
  CheckerState *St = Sp + c;
  const int64_t pass = St->n_pass;
  const uint32_t plo = (uint32_t)pass, phi = (uint32_t)(pass >> 32);
  for (int i = threadIdx.x; i < CHECKER_NTAB; i += SLAB_BLOCK) tab[i] = tables[P.tab_off + i];
  uint8_t *lat = reinterpret_cast<uint8_t *>(lat_pool + P.lat_off);
  const SlabGeom G = slab_geom(P);
  const int klast = k0 + H - 1;
  uint32_t *sm = ring, *s0 = ring + G.pw, *sp = ring + 2 * G.pw;
  slab_load(G, lat, klast - 1, sm, 0, SLAB_BLOCK);
  slab_load(G, lat, klast, s0, 0, SLAB_BLOCK);
  slab_load(G, lat, klast + 1, sp, 0, SLAB_BLOCK);
  __syncthreads();
  SlabDelta d;
  for (int cc = 0; cc < 4; ++cc) {
    slab_colour(G, P, s0, sm, sp, klast, 4 + cc, tab, plo, phi, 0, SLAB_BLOCK, d);
    __syncthreads();
  }
  slab_store(G, lat, klast, s0, 0, SLAB_BLOCK);
  slab_add_deltas(St, d, red);
}

}  // namespace

int checker_slab_ring_bytes(int L0, int L1) {
  return SLAB_RING * (2 * (L0 / 64) + 1) * L1 * 4;
}

bool checker_slab_fits(const CheckerParams &P) {
  return P.ndim == 3 && P.ncolor == 8 && P.L0 % 64 == 0 && P.L1 % 2 == 0 && P.L2 % 2 == 0 &&
         !P.quad_only && checker_slab_ring_bytes(P.L0, P.L1) <= CHECKER_SLAB_MAX_LDS;
}

void launch_checker_slab(const CheckerParams *P, CheckerState *St, int8_t *lat_pool,
                         const uint32_t *tables, const int32_t *slab_work, int n_slabs,
                         int ring_bytes, hipStream_t stream) {
  hipLaunchKernelGGL(checker_slab_kernel, dim3(n_slabs), dim3(SLAB_BLOCK), ring_bytes, stream, P,
                     St, lat_pool, tables, slab_work);
  hipLaunchKernelGGL(checker_slab_seam_kernel, dim3(n_slabs), dim3(SLAB_BLOCK),
                     3 * ring_bytes / SLAB_RING, stream, P, St, lat_pool, tables, slab_work);
}

void *checker_slab_kernel_fn() { return reinterpret_cast<void *>(&checker_slab_kernel); }
void *checker_slab_seam_kernel_fn() { return reinterpret_cast<void *>(&checker_slab_seam_kernel); }
int checker_slab_block() { return SLAB_BLOCK; }
int checker_slab_ring_planes() { return SLAB_RING; }

hipError_t checker_slab_set_max_lds(int bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(checker_slab_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(checker_slab_seam_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace mcg
