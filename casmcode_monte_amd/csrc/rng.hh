// rng.hh -- random number generation shared by host and device code.
//
//  * MT19937-64 (Matsumoto & Nishimura 2000; parameters of the C++ standard's
//    std::mt19937_64, [rand.predef]): the engine of the reference's
//    RandomNumberGenerator (definitions.hh:17).  The state layout (312 words
//    plus the index of the next output) is the one std::mt19937_64 serialises,
//    so a chain's engine can be dumped and loaded like the reference engine.
//  * Lemire's nearly-divisionless bounded integer (arXiv:1805.10941), which is
//    how libstdc++ implements uniform_int_distribution<long>(0, n-1) over a
//    64-bit engine: the reference's site draw (RandomNumberGenerator.hh:31-34).
//  * The canonical double in [0,1) from one 64-bit word, as libstdc++'s
//    generate_canonical<double,53> computes it over a 64-bit engine: round the
//    word to double, divide by 2^64, clamp 1.0 to the largest double below 1
//    (RandomNumberGenerator.hh:38-41).
//  * Philox4x32-10 (Salmon et al., SC'11): the counter-based generator of the
//    checkerboard mode.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define MCG_HD __host__ __device__ inline
#else
#define MCG_HD inline
#endif

namespace mcg {

// --- MT19937-64 parameters ---------------------------------------------------
constexpr int MT_N = 312;
constexpr int MT_M = 156;
constexpr uint64_t MT_A = 0xB5026F5AA96619E9ULL;
constexpr uint64_t MT_UPPER = 0xFFFFFFFF80000000ULL;  // top 33 bits (r = 31)
constexpr uint64_t MT_LOWER = 0x000000007FFFFFFFULL;
constexpr uint64_t MT_F = 6364136223846793005ULL;

MCG_HD uint64_t mt_temper(uint64_t z) {
  z ^= (z >> 29) & 0x5555555555555555ULL;
  z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
  z ^= (z << 37) & 0xFFF7EEE000000000ULL;
  z ^= z >> 43;
  return z;
}

// One word of the twist: new x[k] from x[k] (upper bit), x[k+1] (lower bits)
// and x[k+m] (taken modulo n by the caller).
MCG_HD uint64_t mt_twist_word(uint64_t xk, uint64_t xk1, uint64_t xkm) {
  uint64_t y = (xk & MT_UPPER) | (xk1 & MT_LOWER);
  return xkm ^ (y >> 1) ^ ((y & 1ULL) ? MT_A : 0ULL);
}

// Seeding as std::mt19937_64::seed(value): x[0] = value,
// x[i] = f * (x[i-1] ^ (x[i-1] >> 62)) + i; position = n (twist before the
// first output).
inline void mt_seed(uint64_t seed, uint64_t *x, int *pos) {
  x[0] = seed;
  for (int i = 1; i < MT_N; ++i) x[i] = MT_F * (x[i - 1] ^ (x[i - 1] >> 62)) + (uint64_t)i;
  *pos = MT_N;
}

// Sequential twist of the whole state (host side).
inline void mt_twist(uint64_t *x) {
  for (int k = 0; k < MT_N - MT_M; ++k) x[k] = mt_twist_word(x[k], x[k + 1], x[k + MT_M]);
  for (int k = MT_N - MT_M; k < MT_N - 1; ++k)
    x[k] = mt_twist_word(x[k], x[k + 1], x[k + MT_M - MT_N]);
  x[MT_N - 1] = mt_twist_word(x[MT_N - 1], x[0], x[MT_M - 1]);
}

// Host-side engine with the std::mt19937_64 state layout.
struct Mt64 {
  uint64_t x[MT_N];
  int pos;
  explicit Mt64(uint64_t seed = 5489ULL) { mt_seed(seed, x, &pos); }
  uint64_t operator()() {
    if (pos >= MT_N) {
      mt_twist(x);
      pos = 0;
    }
    return mt_temper(x[pos++]);
  }
};

// --- 64x64 -> 128 multiply ------------------------------------------------------
MCG_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
#endif
}

// Lemire draw of [0, range) given a source of 64-bit words (host reference of
// the device walk): product = x * range; if low < range, redraw while
// low < (2^64 - range) % range; result = high word.
template <typename Src>
inline uint64_t lemire(Src &src, uint64_t range) {
  uint64_t x = src();
  uint64_t lo = x * range;
  uint64_t hi = mulhi64(x, range);
  if (lo < range) {
    uint64_t threshold = (0ULL - range) % range;
    while (lo < threshold) {
      x = src();
      lo = x * range;
      hi = mulhi64(x, range);
    }
  }
  return hi;
}

// libstdc++ canonical double from one 64-bit word.
MCG_HD double canonical_from_word(uint64_t x) {
  double r = (double)x / 18446744073709551616.0;
  if (r >= 1.0) r = 0.99999999999999988898;  // nextafter(1.0, 0.0)
  return r;
}

// --- Philox4x32-10 -------------------------------------------------------------
struct u32x4 {
  uint32_t v[4];
};

MCG_HD void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                         uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  uint64_t p0 = (uint64_t)M0 * c0;
  uint64_t p1 = (uint64_t)M1 * c2;
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c1 ^ k0;
  uint32_t n1 = lo1;
  uint32_t n2 = hi0 ^ c3 ^ k1;
  uint32_t n3 = lo0;
  c0 = n0;
  c1 = n1;
  c2 = n2;
  c3 = n3;
}

MCG_HD u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                           uint32_t k1) {
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c0, c1, c2, c3, k0, k1);
    k0 += W0;
    k1 += W1;
  }
  u32x4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

}  // namespace mcg
