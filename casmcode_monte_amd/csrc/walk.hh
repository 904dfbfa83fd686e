// walk.hh -- wavefront helpers shared by the sequential-chain kernels (EXACT
// single-site flips and canonical pair swaps): scalar broadcasts and the
// out-of-place MT19937-64 twist into a ring of tempered outputs.
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

// Out-of-place twist of generation src into dst (one 64-lane workgroup) and
// temper it into the ring slot.
__device__ inline void twist_gen(const uint64_t *src, uint64_t *dst, uint64_t *slot, int lane) {
  for (int k = lane; k < MT_N - MT_M; k += 64)
    dst[k] = mt_twist_word(src[k], src[k + 1], src[k + MT_M]);
  __syncthreads();
  for (int k = MT_N - MT_M + lane; k < MT_N - 1; k += 64)
    dst[k] = mt_twist_word(src[k], src[k + 1], dst[k - (MT_N - MT_M)]);
  if (lane == 0) dst[MT_N - 1] = mt_twist_word(src[MT_N - 1], dst[0], dst[MT_M - 1]);
  __syncthreads();
  for (int k = lane; k < MT_N; k += 64) slot[k] = mt_temper(dst[k]);
  __syncthreads();
}

}  // namespace mcg
