// fcc.hh -- the binary FCC cluster-expansion stencil shared by the kernels
// (fcc.hip) and the host code that builds tables and counts (api.hip).
//
// Sites are the points (x, y, z), x+y+z even, of a periodic 2L0 x 2L1 x 2L2
// grid (doubled conventional coordinates); site l = h + L0*(y + 2*L1*z) with
// x = ((y+z)&1) + 2h.  A site's interaction range: 12 nearest neighbours
// (squared doubled distance 2), 6 second (4) and 24 third (6) neighbours, and
// the 24 nearest-neighbour triangles through it (pairs of nearest neighbours
// that are nearest neighbours of each other).
#pragma once

#include <cstdint>

// X-macros over the neighbour offsets (k, dx, dy, dz); the order of the 12
// nearest neighbours fixes the bit order of the triangle-count table.
#define MCG_FCC_NN(X)                                                                       \
  X(0, -1, -1, 0) X(1, -1, 0, -1) X(2, -1, 0, 1) X(3, -1, 1, 0) X(4, 0, -1, -1)            \
  X(5, 0, -1, 1) X(6, 0, 1, -1) X(7, 0, 1, 1) X(8, 1, -1, 0) X(9, 1, 0, -1) X(10, 1, 0, 1) \
  X(11, 1, 1, 0)
#define MCG_FCC_N2(X) \
  X(0, -2, 0, 0) X(1, 0, -2, 0) X(2, 0, 0, -2) X(3, 0, 0, 2) X(4, 0, 2, 0) X(5, 2, 0, 0)
#define MCG_FCC_N3(X)                                                                         \
  X(0, -2, -1, -1) X(1, -2, -1, 1) X(2, -2, 1, -1) X(3, -2, 1, 1) X(4, -1, -2, -1)             \
  X(5, -1, -2, 1) X(6, -1, -1, -2) X(7, -1, -1, 2) X(8, -1, 1, -2) X(9, -1, 1, 2)              \
  X(10, -1, 2, -1) X(11, -1, 2, 1) X(12, 1, -2, -1) X(13, 1, -2, 1) X(14, 1, -1, -2)           \
  X(15, 1, -1, 2) X(16, 1, 1, -2) X(17, 1, 1, 2) X(18, 1, 2, -1) X(19, 1, 2, 1)                \
  X(20, 2, -1, -1) X(21, 2, -1, 1) X(22, 2, 1, -1) X(23, 2, 1, 1)

namespace mcg {

constexpr int FCC_Z1 = 12, FCC_Z2 = 6, FCC_Z3 = 24, FCC_ZT = 24;
// acceptance tables per chain (doubles): A[s][n1][n2] then B[s][n3][t],
// s = old occupation; p = A * B = exp(-beta * dE)
constexpr int FCC_TAB_A = (FCC_Z1 + 1) * (FCC_Z2 + 1);   // 91
constexpr int FCC_TAB_B = (FCC_Z3 + 1) * (FCC_ZT + 1);   // 625
constexpr int FCC_NTAB = 2 * (FCC_TAB_A + FCC_TAB_B);    // 1432
constexpr int FCC_LIST = 64;   // longest per-colour coordinate list (2L <= 254)
constexpr int FCC_LMAX = 127;  // conventional cells per edge

// Colour of doubled coordinate v on a ring of n = 2L points: v mod 4, except
// that when n = 2 (mod 4) the last two points get 1 and 2, so that points of
// one colour are 0 or >= 3 apart around the ring.
__host__ __device__ inline int fcc_ring_colour(int v, int n) {
  if ((n & 3) == 0 || v < n - 2) return v & 3;
  return v == n - 2 ? 1 : 2;
}

// LDS line stride of the FCC kernel: L0 sites plus two ghost bytes, plus one
// more when that is a multiple of 4 (lines 4 apart then fall in distinct banks).
__host__ __device__ inline int fcc_lds_stride(int L0) {
  const int s = L0 + 2;
  return (s & 3) == 0 ? s + 1 : s;
}

// Bit-line kernel (fcc.hip, LDS-resident chains with L0 <= 64): one 64-bit
// word per lattice line, stored at ln + ln / 4 -- the line doubled when
// L0 <= FCC_BITS_NARROW_L0, else the line once (doubled in registers);
// dynamic LDS bytes of a chain.
constexpr int FCC_BITS_NARROW_L0 = 32;
constexpr int FCC_BITS_MAX_L0 = 64;
__host__ __device__ inline int fcc_bits_lds_bytes(int L1, int L2) {
  const int n_lines = 4 * L1 * L2;
  return 8 * (n_lines + (n_lines >> 2) + 1);
}

}  // namespace mcg
