// probes.hip -- measured bandwidth peaks of the device (libcasm_monte_probes.so).
//
// Two microbenchmarks give the denominators the benchmark reports beside the
// datasheet figures: a streaming copy through HBM and a read stream from LDS.
// Neither is part of the Metropolis path; bench.py loads this library on its
// own.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/casm_mi355x/mcg_probes.h"

namespace {

thread_local std::string g_err;

int fail(const std::string &m) {
  g_err = m;
  return 1;
}

#define PROBE_TRY(expr)                                                                  \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// 16-byte words, grid-stride: every word read once and written once.
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, int64_t n_words) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words; i += stride)
    dst[i] = src[i];
}

// Each thread reads a 16-byte LDS word per step, cycling over a 16 KB buffer
// with a per-thread stride that keeps a wavefront's 64 reads on distinct
// 16-byte slots (no bank conflicts); the xor of everything read is written
// once per thread so no read can be dropped.
constexpr int LDS_WORDS = 1024;  // 16 KB of uint4
__global__ __launch_bounds__(1024) void lds_read_kernel(int64_t iters, uint32_t *__restrict__ out) {
  __shared__ uint4 buf[LDS_WORDS];
  const int t = threadIdx.x;
  for (int i = t; i < LDS_WORDS; i += blockDim.x)
    buf[i] = make_uint4(i, i * 3u + 1u, i * 5u + 2u, i * 7u + 3u);
  __syncthreads();
  uint32_t acc = 0;
  int idx = t & (LDS_WORDS - 1);
  for (int64_t k = 0; k < iters; ++k) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint4 v = buf[(idx + 64 * u) & (LDS_WORDS - 1)];
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    idx = (idx + 512) & (LDS_WORDS - 1);
  }
  out[(int64_t)blockIdx.x * blockDim.x + t] = acc;
}

}  // namespace

extern "C" {

const char *mcgp_last_error(void) { return g_err.c_str(); }

int mcgp_hbm_copy(int64_t bytes, int64_t reps, double *gbs, double *ms_per_copy) {
  if (bytes < (1 << 20) || (bytes & 15) || reps < 1 || !gbs || !ms_per_copy)
    return fail("mcgp_hbm_copy: bytes must be >= 1 MiB and a multiple of 16, reps >= 1");
  hipDeviceProp_t prop;
  int dev = 0;
  PROBE_TRY(hipGetDevice(&dev));
  PROBE_TRY(hipGetDeviceProperties(&prop, dev));
  uint4 *a = nullptr, *b = nullptr;
  PROBE_TRY(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail("mcgp_hbm_copy: out of device memory");
  }
  hipEvent_t e0, e1;
  int rc = 0;
  float ms = 0.f;
  const int64_t n = bytes / 16;
  const int grid = 8 * prop.multiProcessorCount;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail("mcgp_hbm_copy: hipEventCreate failed");
  } else {
    (void)hipMemset(a, 1, bytes);
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, a, b, n);  // warm-up
    (void)hipEventRecord(e0, 0);
    for (int64_t r = 0; r < reps; ++r)
      hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = fail("mcgp_hbm_copy: launch failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  if (rc) return rc;
  *ms_per_copy = ms / (double)reps;
  *gbs = 2.0 * (double)bytes / (*ms_per_copy * 1e-3) / 1e9;  // read + write
  return 0;
}

int mcgp_lds_read(int64_t iters, double *gbs, double *ms) {
  if (iters < 1 || !gbs || !ms) return fail("mcgp_lds_read: iters >= 1");
  hipDeviceProp_t prop;
  int dev = 0;
  PROBE_TRY(hipGetDevice(&dev));
  PROBE_TRY(hipGetDeviceProperties(&prop, dev));
  // two 1024-thread workgroups per CU (16 KB of LDS each): 32 waves a CU
  const int grid = 2 * prop.multiProcessorCount;
  uint32_t *out = nullptr;
  PROBE_TRY(hipMalloc(&out, sizeof(uint32_t) * (size_t)grid * 1024));
  hipEvent_t e0, e1;
  int rc = 0;
  float t = 0.f;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail("mcgp_lds_read: hipEventCreate failed");
  } else {
    hipLaunchKernelGGL(lds_read_kernel, dim3(grid), dim3(1024), 0, 0, (int64_t)16, out);
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(lds_read_kernel, dim3(grid), dim3(1024), 0, 0, iters, out);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess ||
        hipEventElapsedTime(&t, e0, e1) != hipSuccess)
      rc = fail("mcgp_lds_read: launch failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(out);
  if (rc) return rc;
  *ms = t;
  *gbs = (double)grid * 1024.0 * (double)iters * 8.0 * 16.0 / (t * 1e-3) / 1e9;
  return 0;
}

}  // extern "C"
