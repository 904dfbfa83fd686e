// stats.hip -- BasicStatistics and the equilibration check over a chain's
// device-resident sample series.
//
// Reference: BasicStatisticsCalculator::operator() (src/casm/monte/BasicStatistics.cc:114-131,
// weighted :144-188), autocorrelation_factor (:24-48), resample (:50-73), variance and
// weighted_variance (include/casm/monte/misc/math.hh:31-54), _default_equilibration_check
// (src/casm/monte/checks/EquilibrationCheck.cc:50-117) and its weighted form (:119-162).
//
// The observable of each record is formed on the fly with the reference's FP64
// rounding (the sampling functions of basic_semigrand_canonical.hh:485-590).
// Reductions run in one 256-thread workgroup: the lag search evaluates 256
// consecutive lags at a time (one lag per thread, coalesced reads of the
// centred series from HBM/L2) and stops at the first batch holding a lag with
// |cov/var0| <= 0.5.  Sums are reduced in a different order than the
// reference's sequential loops, so results agree to rounding, not bitwise.
#include <mutex>

#include "internal.hh"

namespace mcg {

// Fixed statistics scratch of the no-allocation entry point
// (mcg_stats_compute_noalloc): the kernels' result words followed by the
// centred or resampled series, MCG_STATS_NOALLOC_MAX doubles in all.  Every GPU
// the library runs on holds its own copy until the process ends.  Host code
// never passes `scratch` itself to a kernel (in host code the name is only a
// stand-in for the symbol); it passes the device address that
// stats_noalloc_scratch_init looks up once per device.
__device__ __align__(256) double scratch[MCG_STATS_NOALLOC_MAX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t stats_noalloc_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev]) return hipSuccess;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch_ptr[dev] = static_cast<double *>(ptr);
  return hipSuccess;
}

double *stats_noalloc_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_ptr[dev];
}

__device__ __forceinline__ double observable_of(const mcg_sample &r, const ObsParams &op) {
  const double N = (double)op.n_sites;
  if (op.fcc) {
    // E_f = N V0 + V1 N1 + V2/6 P1 + V3/3 P2 + V4/12 P3 + Vt/8 T, summed in this order
    double ef = __dmul_rn(op.coef[0], N);
    ef = __dadd_rn(ef, __dmul_rn(op.coef[1], (double)r.M));
    ef = __dadd_rn(ef, __dmul_rn(op.coef[2], (double)r.S1));
    ef = __dadd_rn(ef, __dmul_rn(op.coef[3], (double)r.S2));
    ef = __dadd_rn(ef, __dmul_rn(op.coef[4], (double)r.S3));
    ef = __dadd_rn(ef, __dmul_rn(op.coef[5], (double)r.T3));
    if (op.observable == MCG_OBS_FORMATION_ENERGY) return __ddiv_rn(ef, N);
    const double N1 = (double)r.M;
    if (op.observable == MCG_OBS_PARAM_COMPOSITION) return __ddiv_rn(N1, N);
    return __ddiv_rn(__dsub_rn(ef, __dmul_rn(op.mu, N1)), N);
  }
  // formation energy per supercell
  double ef;
  if (op.fold) {
    ef = r.e_fold;
  } else {
    // fl(S1*(-J1)) [model.hh:266-269], plus the J2 term of the lattice models
    ef = __dmul_rn((double)r.S1, op.negJ1);
    if (op.negJ2 != 0.0) ef = __dadd_rn(ef, __dmul_rn((double)r.S2, op.negJ2));
  }
  if (op.observable == MCG_OBS_FORMATION_ENERGY) return __ddiv_rn(ef, N);
  // OCCUPANT: the potential has an exchange potential per species, carried in
  // the running energy per unit cell
  if (op.running && op.observable == MCG_OBS_POTENTIAL_ENERGY) return r.e_run;
  const double Nx = __ddiv_rn((double)(op.n_sites + r.M), 2.0);  // model.hh:415
  if (op.observable == MCG_OBS_PARAM_COMPOSITION) return __ddiv_rn(Nx, N);
  // basic_semigrand_canonical.hh:165-174
  const double e = __dsub_rn(ef, __dmul_rn(op.mu, Nx));
  return __ddiv_rn(e, N);
}

template <int BS>
__device__ double block_sum(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = __dadd_rn(red[t], red[t + s]);
    __syncthreads();
  }
  double out = red[0];
  __syncthreads();
  return out;
}

// autocorrelation_factor (BasicStatistics.cc:24-48) of a centred series
// d[0..count) with variance var0 and mean `mean`; `increment` is the interval
// between resampled observations (1 for unweighted ones).  The whole
// workgroup takes part; the lag found is returned in *lag_out (0 if none).
__device__ double lag_factor(const double *__restrict__ d, int64_t count, double var0,
                             double mean, double increment, int *best, double *lag_out) {
  const int t = threadIdx.x;
  double lag = 0.0;
  double factor;
  if (fabs(var0 / mean) < 1e-8 || var0 == 0.0) {
    factor = 1.0;
  } else {
    factor = 1.7976931348623157e308;  // DBL_MAX: no lag found
    for (int64_t base = 1; base < count; base += 256) {
      if (t == 0) *best = 0x7fffffff;
      __syncthreads();
      const int64_t i = base + t;
      if (i < count) {
        const int64_t range = count - i;
        double cov = 0.0;
        for (int64_t j = 0; j < range; ++j) cov = __dadd_rn(cov, __dmul_rn(d[j], d[j + i]));
        cov = __ddiv_rn(cov, (double)range);
        if (fabs(cov / var0) <= 0.5) atomicMin(best, t);
      }
      __syncthreads();
      if (*best != 0x7fffffff) {
        lag = (double)(base + *best);
        break;
      }
      __syncthreads();
    }
    if (lag > 0.0) {
      const double rho = pow(2.0, -1.0 / (lag * increment));
      factor = (1.0 + rho) / (1.0 - rho);
    }
  }
  *lag_out = lag;
  return factor;
}

// out[0] = mean, out[1] = var0, out[2] = autocorrelation factor, out[3] = lag
// d_centered: scratch of `count` doubles; the whole 256-thread block takes part
__device__ void block_stats(const mcg_sample *__restrict__ series, int64_t first, int64_t count,
                            const ObsParams &op, double *__restrict__ d_centered,
                            double *__restrict__ out, double *red, int *best_p) {
  int &best = *best_p;
  const int t = threadIdx.x;
  double s = 0.0;
  for (int64_t j = t; j < count; j += 256) s = __dadd_rn(s, observable_of(series[first + j], op));
  const double mean = __ddiv_rn(block_sum<256>(s, red), (double)count);
  double v = 0.0;
  for (int64_t j = t; j < count; j += 256) {
    const double d = __dsub_rn(observable_of(series[first + j], op), mean);
    d_centered[j] = d;
    v = __dadd_rn(v, __dmul_rn(d, d));
  }
  const double var0 = __ddiv_rn(block_sum<256>(v, red), (double)count);
  __syncthreads();
  double lag;
  const double factor = lag_factor(d_centered, count, var0, mean, 1.0, &best, &lag);
  if (t == 0) {
    out[0] = mean;
    out[1] = var0;
    out[2] = factor;
    out[3] = lag;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void stats_kernel(const mcg_sample *__restrict__ series,
                                                    int64_t first, int64_t count, ObsParams op,
                                                    double *__restrict__ d_centered,
                                                    double *__restrict__ out) {
  __shared__ double red[256];
  __shared__ int best;
  block_stats(series, first, count, op, d_centered, out, red, &best);
}

// Weighted observations (the records' weights), BasicStatistics.cc:144-188:
// the series read at n equally spaced points of its cumulative weight
// (resample, :50-73, one lane in order), then method 1 (weighted mean and
// variance, misc/math.hh:45-54; autocorrelation of the resampled series with
// increment W/n) or method 2 (unweighted statistics of the resampled series).
// out[0] = mean, out[1] = variance, out[2] = autocorrelation factor,
// out[3] = lag, out[4] = W (the sum the precision divides by: W or n);
// d_eq: scratch of n doubles
__global__ __launch_bounds__(256) void weighted_stats_kernel(
    const mcg_sample *__restrict__ series, int64_t first, int64_t count, ObsParams op,
    int method, int64_t n, double *__restrict__ d_eq, double *__restrict__ out) {
  __shared__ double red[256];
  __shared__ int best;
  const int t = threadIdx.x;
  const mcg_sample *S = series + first;
  double sw = 0.0, sxw = 0.0;
  for (int64_t j = t; j < count; j += 256) {
    const double w = S[j].weight;
    sw = __dadd_rn(sw, w);
    sxw = __dadd_rn(sxw, __dmul_rn(observable_of(S[j], op), w));
  }
  const double W = block_sum<256>(sw, red);
  const double XW = block_sum<256>(sxw, red);
  const double increment = __ddiv_rn(W, (double)n);
  if (t == 0) {
    int64_t j = 0;
    double Wj = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double target = __dmul_rn((double)i, increment);
      while (j + 1 < count && __dadd_rn(Wj, S[j].weight) < target) {
        Wj = __dadd_rn(Wj, S[j].weight);
        ++j;
      }
      d_eq[i] = observable_of(S[j], op);
    }
  }
  __syncthreads();
  double se = 0.0;
  for (int64_t i = t; i < n; i += 256) se = __dadd_rn(se, d_eq[i]);
  const double eq_mean = __ddiv_rn(block_sum<256>(se, red), (double)n);
  double ve = 0.0;
  for (int64_t i = t; i < n; i += 256) {
    const double d = __dsub_rn(d_eq[i], eq_mean);
    d_eq[i] = d;
    ve = __dadd_rn(ve, __dmul_rn(d, d));
  }
  const double eq_var = __ddiv_rn(block_sum<256>(ve, red), (double)n);
  __syncthreads();
  double mean, var, denom, lag, factor;
  if (method == 1) {
    mean = __ddiv_rn(XW, W);
    double sv = 0.0;
    for (int64_t j = t; j < count; j += 256) {
      const double d = __dsub_rn(observable_of(S[j], op), mean);
      sv = __dadd_rn(sv, __dmul_rn(__dmul_rn(S[j].weight, d), d));  // w(i) * d * d
    }
    var = __ddiv_rn(block_sum<256>(sv, red), W);
    denom = W;
    factor = lag_factor(d_eq, n, eq_var, eq_mean, increment, &best, &lag);
  } else {
    mean = eq_mean;
    var = eq_var;
    denom = (double)n;
    factor = lag_factor(d_eq, n, eq_var, eq_mean, 1.0, &best, &lag);
  }
  if (t == 0) {
    out[0] = mean;
    out[1] = var;
    out[2] = factor;
    out[3] = lag;
    out[4] = denom;
  }
}

// Sequential scan of _default_equilibration_check by one lane.  With weights
// (default_equilibration_check, EquilibrationCheck.cc:134-160) each observation
// is scaled by (N/W) * w_i first, W the sum of the weights in order.
// out[0] = is_equilibrated, out[1] = N_samples_for_equilibration
__device__ void equilibration_scan(const mcg_sample *__restrict__ series, int64_t count,
                                   const ObsParams &op, double prec, int weighted,
                                   int64_t *__restrict__ out) {
  double wf = 0.0;
  if (weighted) {
    double W = 0.0;
    for (int64_t i = 0; i < count; ++i) W = __dadd_rn(W, series[i].weight);
    wf = __ddiv_rn((double)count, W);
  }
  auto obs = [&](int64_t i) {
    const double x = observable_of(series[i], op);
    return weighted ? __dmul_rn(x, __dmul_rn(wf, series[i].weight)) : x;
  };
  const int64_t N = count;
  const double x0 = obs(0);
  const double eps = (x0 == 0.0) ? 1e-8 : fabs(x0) * 1e-8;
  bool all_same = true;
  for (int64_t i = 0; i < N; ++i)
    if (fabs(obs(i) - x0) > eps) {
      all_same = false;
      break;
    }
  if (all_same) {
    out[0] = 1;
    out[1] = 0;
    return;
  }
  bool is_even = (N % 2) == 0;
  int64_t start1 = 0;
  int64_t start2 = is_even ? N / 2 : N / 2 + 1;
  double sum1 = 0.0, sum2 = 0.0;
  for (int64_t i = 0; i < start2; ++i) sum1 = __dadd_rn(sum1, obs(i));
  for (int64_t i = start2; i < N; ++i) sum2 = __dadd_rn(sum2, obs(i));
  while (fabs(sum1 / (double)(start2 - start1) - sum2 / (double)(N - start2)) > prec &&
         start1 < N - 2) {
    if (is_even) {
      sum1 = __dsub_rn(sum1, obs(start1));
      sum1 = __dadd_rn(sum1, obs(start2));
      sum2 = __dsub_rn(sum2, obs(start2));
      start2++;
    } else {
      sum1 = __dsub_rn(sum1, obs(start1));
    }
    start1++;
    is_even = !is_even;
  }
  const double mean_tot = (sum1 + sum2) / (double)(N - start1);
  if (obs(start1) < mean_tot) {
    while (obs(start1) < mean_tot && start1 < N - 1) start1++;
  } else {
    while (obs(start1) > mean_tot && start1 < N - 1) start1++;
  }
  out[0] = (start1 < N - 1) ? 1 : 0;
  out[1] = start1;
}

__global__ void equilibration_kernel(const mcg_sample *__restrict__ series, int64_t count,
                                     ObsParams op, double prec, int weighted,
                                     int64_t *__restrict__ out) {
  if (threadIdx.x != 0) return;
  equilibration_scan(series, count, op, prec, weighted, out);
}

// The completion check of a list of chains at once (one block per chain): the
// equilibration scan of the whole series at precision prec (one lane), then,
// when equilibrated with samples left, block_stats over the samples after the
// equilibration index.  out: 8 doubles per chain -- is_equilibrated,
// n_equilibration, mean, var0, autocorrelation factor, lag, count of the tail
// (0 when no statistics were computed); scratch: the centred tails at
// scratch_off[b].
__global__ __launch_bounds__(256) void check_batch_kernel(
    const mcg_sample *__restrict__ series, int64_t series_cap,
    const int64_t *__restrict__ chains, const int64_t *__restrict__ counts,
    const ObsParams *__restrict__ ops, const int64_t *__restrict__ scratch_off, double prec,
    double *__restrict__ scratch, double *__restrict__ out) {
  __shared__ double red[256];
  __shared__ int best;
  __shared__ int64_t eq[2];
  __shared__ double res[4];
  const int b = blockIdx.x;
  const int64_t count = counts[b];
  const mcg_sample *S = series + chains[b] * series_cap;
  const ObsParams op = ops[b];
  if (threadIdx.x == 0) equilibration_scan(S, count, op, prec, 0, eq);
  __syncthreads();
  const int64_t is_eq = eq[0], n_eq = eq[1];
  double *o = out + 8 * (int64_t)b;
  if (is_eq && n_eq < count) {
    block_stats(S, n_eq, count - n_eq, op, scratch + scratch_off[b], res, red, &best);
    if (threadIdx.x == 0) {
      o[2] = res[0];
      o[3] = res[1];
      o[4] = res[2];
      o[5] = res[3];
      o[6] = (double)(count - n_eq);
    }
  } else if (threadIdx.x == 0) {
    o[2] = o[3] = o[4] = o[5] = o[6] = 0.0;
  }
  if (threadIdx.x == 0) {
    o[0] = (double)is_eq;
    o[1] = (double)n_eq;
  }
}

void launch_stats(const mcg_sample *series, int64_t first, int64_t count, ObsParams op,
                  double *d_out, hipStream_t stream) {
  // d_out: 4 result doubles followed by `count` doubles of scratch
  hipLaunchKernelGGL(stats_kernel, dim3(1), dim3(256), 0, stream, series, first, count, op,
                     d_out + 4, d_out);
}

void launch_weighted_stats(const mcg_sample *series, int64_t first, int64_t count, ObsParams op,
                           int method, int64_t n_resamples, double *d_out, hipStream_t stream) {
  // d_out: 8 result doubles followed by n_resamples doubles of scratch
  hipLaunchKernelGGL(weighted_stats_kernel, dim3(1), dim3(256), 0, stream, series, first, count,
                     op, method, n_resamples, d_out + 8, d_out);
}

void launch_check_batch(const mcg_sample *series, int64_t series_cap, const int64_t *chains,
                        const int64_t *counts, const ObsParams *ops, const int64_t *scratch_off,
                        int n, double prec, double *scratch, double *out, hipStream_t stream) {
  hipLaunchKernelGGL(check_batch_kernel, dim3(n), dim3(256), 0, stream, series, series_cap,
                     chains, counts, ops, scratch_off, prec, scratch, out);
}

void launch_equilibration(const mcg_sample *series, int64_t count, ObsParams op, double prec,
                          int weighted, int64_t *d_out, hipStream_t stream) {
  hipLaunchKernelGGL(equilibration_kernel, dim3(1), dim3(64), 0, stream, series, count, op,
                     prec, weighted, d_out);
}

}  // namespace mcg
