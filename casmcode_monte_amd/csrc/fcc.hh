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
// that when n = 2 (mod 4) the last two points get colours of their own, 4
// and 5, so that points of one colour are 0 or >= 3 apart around the ring and
// colours 0-3 keep one parity each (as on rings of n = 0 mod 4, no two points
// of colours 0-3 within distance 2 of each other straddle the seam).
constexpr int FCC_NCOL = 6;
constexpr int FCC_NCLASS = FCC_NCOL * FCC_NCOL * FCC_NCOL;  // col = cx + 6 cy + 36 cz
__host__ __device__ inline int fcc_ring_colour(int v, int n) {
  if ((n & 3) == 0 || v < n - 2) return v & 3;
  return v == n - 2 ? 4 : 5;
}

// LDS line stride of the FCC kernel: L0 sites plus two ghost bytes, plus one
// more when that is a multiple of 4 (lines 4 apart then fall in distinct banks).
__host__ __device__ inline int fcc_lds_stride(int L0) {
  const int s = L0 + 2;
  return (s & 3) == 0 ? s + 1 : s;
}

// The 24 nearest-neighbour triangles through a site as pairs of bits of the
// nearest-neighbour mask of the bit-line kernels (bit i: offset (dx, dy, dz) =
// (-1,-1,0), (1,-1,0), (-1,1,0), (1,1,0), (-1,0,-1), (1,0,-1), (-1,0,1),
// (1,0,1), (0,-1,-1), (0,-1,1), (0,1,-1), (0,1,1)): the pairs whose offsets
// differ by a nearest-neighbour offset.  A handle checks them against the
// host's triangle table in that bit order when it is built.
constexpr uint8_t FCC_TRI_PAIRS[24][2] = {
    {0, 4}, {0, 6}, {0, 8}, {0, 9}, {1, 5}, {1, 7}, {1, 8}, {1, 9},
    {2, 4}, {2, 6}, {2, 10}, {2, 11}, {3, 5}, {3, 7}, {3, 10}, {3, 11},
    {4, 8}, {4, 10}, {5, 8}, {5, 10}, {6, 9}, {6, 11}, {7, 9}, {7, 11}};

// Class schedule of a pass: the non-empty colour classes in the order they
// are swept, grouped into steps of classes that do not interact (no site of
// one lies within the interaction range of a site of another), so the
// classes of a step update together, with one barrier per step.  Built on the
// host for each chain's shape (fcc_class_schedule in api.hip): classes in
// order of decreasing size (then class id), each joining the first step none
// of whose classes it interacts with (at most FCC_STEP_MAX per step).  Every
// FCC kernel sweeps this schedule, so all give one trajectory: 8 steps of 4
// classes when every edge is even, 20 (8 large, 12 on the odd rings' seams)
// when every edge is odd.
constexpr int FCC_STEP_MAX = 8;
constexpr int FCC_SCHED_MAX = 128;  // non-empty classes: 32 to 108
struct FccSched {
  int32_t n_steps;
  int32_t n_classes;
  uint8_t start[FCC_SCHED_MAX + 1];  // step s holds cls[start[s] .. start[s+1])
  uint8_t cls[FCC_SCHED_MAX];        // classes (col = cx + 6 cy + 36 cz) in sweep order
  uint8_t pad[7];
};

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
