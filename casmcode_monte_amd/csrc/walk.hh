// walk.hh -- wavefront helpers shared by the sequential-chain kernels (EXACT
// single-site flips, canonical pair swaps, N-fold way): scalar broadcasts, the
// out-of-place MT19937-64 twist into a ring of tempered outputs, and the
// producer/consumer hand-off through which a second wavefront twists the
// engine while the walking wavefront consumes it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rng.hh"

namespace mcg {

// Two generations of tempered outputs: the walk reads stream positions
// [312(G-1), 312(G+1)) while generation G is the newest.
constexpr int RING = 2 * MT_N;

__device__ __forceinline__ uint64_t rfl64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int32_t rfl32(int32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  uint64_t u = __double_as_longlong(v);
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
  uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Out-of-place twist of generation src into dst by one 64-lane wavefront
// (the workgroup), tempering the new words into the ring slot.  Lane L owns
// words L, L+64, L+128 (the first 156, which read only old words) and L+156,
// L+220, L+284 (the rest, each of which reads the new word 156 below it: one
// this lane computed).  Word 311 also reads the new word 0 (lane 0's).  All
// reads go out before any arithmetic, and since src and dst differ no write can
// disturb a read: one LDS round trip per generation.  The wave's LDS operations
// execute in issue order, so the caller's later reads see these writes.
__device__ inline void twist_gen(const uint64_t *src, uint64_t *dst, uint64_t *slot, int lane) {
  const int L = lane;
  const bool third = L < MT_N - MT_M - 128;  // words L+128 and L+284 exist (L < 28)
  const bool wraps = L == MT_N - MT_M - 128 - 1;  // L+284 == 311
  const uint64_t a0 = src[L], a0n = src[L + 1], a1 = src[L + 64], a1n = src[L + 65];
  const uint64_t b0 = src[L + 156], b0n = src[L + 157], b1 = src[L + 220], b1n = src[L + 221];
  uint64_t a2 = 0, a2n = 0, b2 = 0, b2n = 0;
  if (third) {
    a2 = src[L + 128];
    a2n = src[L + 129];
    b2 = src[L + 284];
    if (!wraps) b2n = src[L + 285];
  }
  const uint64_t d0 = mt_twist_word(a0, a0n, b0);
  const uint64_t d1 = mt_twist_word(a1, a1n, b1);
  const uint64_t d2 = mt_twist_word(a2, a2n, b2);
  const uint64_t new0 = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(d0 >> 32), 0) << 32) |
                        (uint32_t)__builtin_amdgcn_readlane((uint32_t)d0, 0);
  const uint64_t e0 = mt_twist_word(b0, b0n, d0);
  const uint64_t e1 = mt_twist_word(b1, b1n, d1);
  const uint64_t e2 = mt_twist_word(b2, wraps ? new0 : b2n, d2);
  dst[L] = d0;
  dst[L + 64] = d1;
  dst[L + 156] = e0;
  dst[L + 220] = e1;
  slot[L] = mt_temper(d0);
  slot[L + 64] = mt_temper(d1);
  slot[L + 156] = mt_temper(e0);
  slot[L + 220] = mt_temper(e1);
  if (third) {
    dst[L + 128] = d2;
    dst[L + 284] = e2;
    slot[L + 128] = mt_temper(d2);
    slot[L + 284] = mt_temper(e2);
  }
  asm volatile("" ::: "memory");
}

// The engine's generations come from a second wavefront of the workgroup.
// raw[g % 3] holds generation g untempered and ring[g % 3] its tempered
// outputs.  The walk reads two generations at a time (G-1 and G, G the newest
// it has), so the producer can write generation G+1 into the third slot while
// the walk works.  The waves meet at one workgroup barrier per generation:
//   B_g: generation g+1 is in its slots; the walk passes it when its window
//   first needs generation g+1, the producer after writing it.
// The producer twists generation g right after B_{g-2}: by then the walk's
// window reaches into generation g-1, so its last consumed word lies in
// generation g-2 or later and the slots of generation g-3 are free.  Both waves
// count the barriers.  When the walk is done it writes the number its final
// barrier will have into `done` and passes that barrier, which is the one the
// producer waits at; the producer returns after the barrier whose number equals
// `done` (a read that races ahead of the final barrier sees a larger number and
// goes on to the final barrier).
__device__ inline void mt_producer(uint64_t (*raw)[MT_N], uint64_t (*ring)[MT_N],
                            volatile const int *done, int L) {
  const bool third = L < MT_N - MT_M - 128;
  const bool wraps = L == MT_N - MT_M - 128 - 1;
  int nbar = 0;
  for (int64_t g = 1;; ++g) {
    const uint64_t *src = raw[(g - 1) % 3];
    uint64_t *dst = raw[g % 3];
    uint64_t *slot = ring[g % 3];
    // generation g exactly as twist_gen forms it (lane L: words L, L+64,
    // L+128, L+156, L+220, L+284)
    const uint64_t a0 = src[L], a0n = src[L + 1], a1 = src[L + 64], a1n = src[L + 65];
    const uint64_t b0 = src[L + 156], b0n = src[L + 157], b1 = src[L + 220], b1n = src[L + 221];
    uint64_t a2 = 0, a2n = 0, b2 = 0, b2n = 0;
    if (third) {
      a2 = src[L + 128];
      a2n = src[L + 129];
      b2 = src[L + 284];
      if (!wraps) b2n = src[L + 285];
    }
    const uint64_t d0 = mt_twist_word(a0, a0n, b0);
    const uint64_t d1 = mt_twist_word(a1, a1n, b1);
    const uint64_t d2 = mt_twist_word(a2, a2n, b2);
    const uint64_t new0 = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(d0 >> 32), 0) << 32) |
                          (uint32_t)__builtin_amdgcn_readlane((uint32_t)d0, 0);
    const uint64_t e0 = mt_twist_word(b0, b0n, d0);
    const uint64_t e1 = mt_twist_word(b1, b1n, d1);
    const uint64_t e2 = mt_twist_word(b2, wraps ? new0 : b2n, d2);
    dst[L] = d0;
    dst[L + 64] = d1;
    dst[L + 156] = e0;
    dst[L + 220] = e1;
    slot[L] = mt_temper(d0);
    slot[L + 64] = mt_temper(d1);
    slot[L + 156] = mt_temper(e0);
    slot[L + 220] = mt_temper(e1);
    if (third) {
      dst[L + 128] = d2;
      dst[L + 284] = e2;
      slot[L + 128] = mt_temper(d2);
      slot[L + 284] = mt_temper(e2);
    }
    __syncthreads();  // B_{g-1}
    if (*done == ++nbar) return;
  }
}

// The walking wave's side of the hand-off: G is the newest generation it may
// read (the slots hold G-1 and G), nbar the barriers passed.
struct MtRing3 {
  int64_t G = 0;
  int sG = 0;  // G % 3
  int nbar = 0;
  // make stream positions < upto readable
  __device__ __forceinline__ void ensure(int64_t upto) {
    while (upto > (int64_t)MT_N * (G + 1)) {
      __syncthreads();  // B_G: generation G+1 is in its slots
      ++nbar;
      ++G;
      sG = sG == 2 ? 0 : sG + 1;
    }
  }
  // offset of stream position q, 312(G-1) <= q < 312(G+1), in ring[3][MT_N]:
  // generation G sits in slot sG, generation G-1 in slot (sG + 2) % 3, so a
  // negative r lands at r + 312 sG except when sG = 0 (slot 2: r + 936); one
  // select, no branch
  __device__ __forceinline__ int index(int64_t q) const {
    const int r = (int)(q - (int64_t)MT_N * G);
    return r + MT_N * sG + ((r < 0 && sG == 0) ? 3 * MT_N : 0);
  }
  // the walk is done: release the producer (lane 0 writes the final barrier's
  // number; every lane of the walking wave passes the barrier)
  __device__ __forceinline__ void finish(int *done, int lane) {
    if (lane == 0) *done = nbar + 1;
    __syncthreads();
  }
};

// Starts of the attempts in a window of W stream words (bit t: an attempt
// starts at word t) when an attempt takes two words or, if it draws (bit t of
// ND), three: 0, then 2 or 3 words on, up to the first start in SX (excluded:
// the walk stops before it) or in SI (included: the walk stops after it).
// Starts keep their parity until one draws, so they are taken a segment at a
// time, one iteration per drawing attempt instead of one per attempt.
__device__ __forceinline__ uint64_t window_starts2(uint64_t ND, uint64_t SX, uint64_t SI, int W) {
  const uint64_t win = W >= 64 ? ~0ull : ((1ull << W) - 1ull);
  uint64_t A = 0;
  int t = 0;
  while (t < W) {
    const uint64_t res = (t & 1) ? 0xAAAAAAAAAAAAAAAAull : 0x5555555555555555ull;
    const uint64_t cand = res & (~0ull << t) & win;
    const uint64_t stop = cand & (SX | SI | ND);
    if (!stop) {
      A |= cand;
      break;
    }
    const int f = __builtin_ctzll(stop);
    const uint64_t below = (1ull << f) - 1ull;
    if ((SX >> f) & 1ull) {
      A |= cand & below;
      break;
    }
    A |= cand & (below | (1ull << f));
    if ((SI >> f) & 1ull) break;
    t = f + 3;
  }
  return A;
}

}  // namespace mcg
