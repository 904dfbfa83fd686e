// refcpu.hh -- CPU restatement of libcasm-monte's occupation-Metropolis hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is compiled into
// oracle/lib/liboracle.so and is used by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg, and nowhere else.  The product (the HIP library
// under casmcode_monte_amd/) never includes, links or calls it.
//
// It restates, in plain C++17 over the host libstdc++ <random> and glibc exp,
// the reference algorithm as read from these files (paths relative to the
// reference repository root):
//   include/casm/monte/ising_cpp/model.hh                 (configuration, energies)
//   include/casm/monte/ising_cpp/basic_semigrand_canonical.hh (SGC potential, events)
//   include/casm/monte/methods/metropolis.hh              (acceptance)
//   include/casm/monte/methods/basic_occupation_metropolis.hh (main loop)
//   include/casm/monte/methods/occupation_metropolis.hh   (RunManager loop)
//   include/casm/monte/RandomNumberGenerator.hh, definitions.hh (RNG)
//   include/casm/monte/sampling/Sampler.hh, SamplingParams.hh
//   include/casm/monte/checks/{CompletionCheck,CutoffCheck,ConvergenceCheck}.hh
//   src/casm/monte/checks/EquilibrationCheck.cc
//   src/casm/monte/BasicStatistics.cc, include/casm/monte/misc/math.hh
//   include/casm/monte/run_management/{RunManager,SamplingFixture}.hh
//   src/casm/monte/events/{OccLocation,OccCandidate}.cc,
//   include/casm/monte/events/{OccLocation,OccEventProposal}.hh (canonical swaps)
// Each function cites the file:line it follows.
//
// Models that the reference does not define (2D with next-nearest J2, 3D simple
// cubic, binary FCC cluster expansion) are defined here in the reference's
// conventions (E = -J * sum over bonds of s_i s_j; SGC potential E - mu * N * x).
// Their pinning is by exact enumeration of tiny lattices, closed forms and
// hand-counted ordered structures, in tests/test_oracle_kats.py and
// tests/test_oracle_fcc.py.
#pragma once

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <optional>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace refcpu {

// Boltzmann constant in eV/K as compiled into CASMcode_global (value pinned by
// this project; CASMcode_global is not available to read).
constexpr double KB = 8.6173303e-05;

typedef long CountType;        // definitions.hh:25
typedef long long BigCountType;  // definitions.hh:26
typedef long Index;

// ---------------------------------------------------------------------------
// RNG: RandomNumberGenerator.hh:14-42 over std::mt19937_64 (definitions.hh:17)
// ---------------------------------------------------------------------------
struct RandomNumberGenerator {
  std::shared_ptr<std::mt19937_64> engine;
  explicit RandomNumberGenerator(std::shared_ptr<std::mt19937_64> e)
      : engine(std::move(e)) {
    if (!engine) throw std::runtime_error("refcpu: null engine");
  }
  // RandomNumberGenerator.hh:31-34
  long random_int(long maximum_value) {
    return std::uniform_int_distribution<long>(0, maximum_value)(*engine);
  }
  // RandomNumberGenerator.hh:38-41
  double random_real(double maximum_value) {
    return std::uniform_real_distribution<double>(0, maximum_value)(*engine);
  }
};

// metropolis.hh:25-35
template <typename G>
bool metropolis_acceptance(double delta_potential_energy, double beta, G &rng) {
  if (delta_potential_energy < 0.0) return true;
  double rand = rng.random_real(1.0);
  double prob = std::exp(-delta_potential_energy * beta);
  return rand < prob;
}

// SamplingParams.hh:193-204: counts to the next sample of a stochastic
// sampling period -- one uniform draw on [0, 1) per count until a draw falls
// below the rate
template <typename G>
CountType stochastic_count_step(double sample_rate, G &rng) {
  CountType dn = 1;
  while (true) {
    if (rng.random_real(1.0) < sample_rate) return dn;
    ++dn;
  }
}

// ---------------------------------------------------------------------------
// Ising model: model.hh:19-436
// ---------------------------------------------------------------------------
struct IsingConfiguration {  // model.hh:19-110
  Index rows = 0, cols = 0;  // shape[0], shape[1]
  Index n_sites = 0;
  std::vector<int> occupation;  // column-major: l = rows*col + row
  IsingConfiguration() = default;
  IsingConfiguration(Index r, Index c, int fill = 1)
      : rows(r), cols(c), n_sites(r * c), occupation(r * c, fill) {}
  int occ(Index l) const { return occupation[l]; }
  void set_occ(Index l, int v) { occupation[l] = v; }
  Index within(Index index, int dim) const {  // model.hh:73-79
    Index s = dim == 0 ? rows : cols;
    Index r = index % s;
    if (r < 0) r += s;
    return r;
  }
  Index to_linear(Index i, Index j) const { return rows * j + i; }  // :93-99
};

struct IsingFormationEnergy {  // model.hh:164-380
  double J = 1.0;
  bool use_nlist = true;
  IsingConfiguration const *config = nullptr;
  std::vector<std::vector<Index>> nlist, flower_nlist;

  IsingFormationEnergy(double _J, bool _use_nlist) : J(_J), use_nlist(_use_nlist) {}

  void set_state(IsingConfiguration const *c) {  // model.hh:210-256
    config = c;
    if (!use_nlist) return;
    nlist.assign(c->n_sites, {});
    flower_nlist.assign(c->n_sites, {});
    for (Index l = 0; l < c->n_sites; ++l) {
      Index i = l % c->rows, j = l / c->rows;
      Index in = c->within(i + 1, 0);
      nlist[l].push_back(c->to_linear(in, j));
      flower_nlist[l].push_back(c->to_linear(in, j));
      Index jn = c->within(j + 1, 1);
      nlist[l].push_back(c->to_linear(i, jn));
      flower_nlist[l].push_back(c->to_linear(i, jn));
      in = c->within(i - 1, 0);
      flower_nlist[l].push_back(c->to_linear(in, j));
      jn = c->within(j - 1, 1);
      flower_nlist[l].push_back(c->to_linear(i, jn));
    }
  }

  double per_supercell() const {  // model.hh:259-290
    auto const &occ = config->occupation;
    if (use_nlist) {
      double e = 0.0;
      for (Index l = 0; l < config->n_sites; ++l)
        e += occ[l] * (occ[nlist[l][0]] + occ[nlist[l][1]]);
      e *= -J;
      return e;
    }
    Index rows = config->rows, cols = config->cols;
    auto site = [&](Index i, Index j) { return occ[i + rows * j]; };
    double e = 0.0;
    for (Index i = 0; i < rows; ++i) {  // row(i) . row(i+1): integer dot
      Index in = config->within(i + 1, 0);
      long dot = 0;
      for (Index j = 0; j < cols; ++j) dot += site(i, j) * site(in, j);
      e += -J * static_cast<double>(dot);
    }
    for (Index j = 0; j < cols; ++j) {  // col(j) . col(j+1)
      Index jn = config->within(j + 1, 1);
      long dot = 0;
      for (Index i = 0; i < rows; ++i) dot += site(i, j) * site(i, jn);
      e += -J * static_cast<double>(dot);
    }
    return e;
  }
  double per_unitcell() const {  // model.hh:293-295
    return per_supercell() / config->n_sites;
  }
  double single_delta(Index l, int new_occ) const {  // model.hh:305-345
    auto const &occ = config->occupation;
    if (use_nlist) {
      return -J * (new_occ - occ[l]) *
             (occ[flower_nlist[l][0]] + occ[flower_nlist[l][1]] +
              occ[flower_nlist[l][2]] + occ[flower_nlist[l][3]]);
    }
    Index rows = config->rows;
    Index i = l % rows, j = l / rows;
    auto site = [&](Index a, Index b) { return occ[a + rows * b]; };
    double ds = new_occ - site(i, j);
    return -J * ds *
           (site(i, config->within(j - 1, 1)) + site(i, config->within(j + 1, 1)) +
            site(config->within(i - 1, 0), j) + site(config->within(i + 1, 0), j));
  }
};

struct IsingParamComposition {  // model.hh:388-436
  IsingConfiguration const *config = nullptr;
  double per_supercell() const {  // :412-417
    long sum = 0;
    for (int v : config->occupation) sum += v;
    return static_cast<double>(config->n_sites + sum) / 2.0;
  }
  double per_unitcell() const { return per_supercell() / config->n_sites; }  // :420-422
  double delta(Index l, int new_occ) const {  // :425-435
    double Ndx = 0.0;
    Ndx += (new_occ - config->occ(l)) / 2.0;
    return Ndx;
  }
};

// basic_semigrand_canonical.hh:104-199 (one exchange potential component)
struct SemiGrandCanonicalPotential {
  IsingFormationEnergy formation_energy;
  IsingParamComposition param_composition;
  double mu = 0.0;
  SemiGrandCanonicalPotential(IsingFormationEnergy fe) : formation_energy(std::move(fe)) {}
  void set_state(IsingConfiguration const *c, double _mu) {
    formation_energy.set_state(c);
    param_composition.config = c;
    mu = _mu;
  }
  double per_supercell() const {  // :165-169
    return formation_energy.per_supercell() - mu * param_composition.per_supercell();
  }
  double per_unitcell() const {  // :172-174
    return per_supercell() / formation_energy.config->n_sites;
  }
  double delta(Index l, int new_occ) const {  // :178-192
    double dEf = formation_energy.single_delta(l, new_occ);
    double Ndx = param_composition.delta(l, new_occ);
    return dEf - mu * Ndx;
  }
};

// ---------------------------------------------------------------------------
// Sampling: Sampler.hh:30-112, 128-180
// ---------------------------------------------------------------------------
struct Sampler {
  Index n_components = 1;
  std::vector<std::vector<double>> rows;  // one row per sample
  explicit Sampler(Index nc = 1) : n_components(nc) {}
  void push_back(std::vector<double> const &v) { rows.push_back(v); }
  CountType n_samples() const { return rows.size(); }
  std::vector<double> component(Index c) const {
    std::vector<double> out(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) out[i] = rows[i][c];
    return out;
  }
};
typedef std::map<std::string, std::shared_ptr<Sampler>> SamplerMap;

inline CountType get_n_samples(SamplerMap const &s) {  // Sampler.hh:410-416
  if (s.size()) return s.begin()->second->n_samples();
  return 0;
}

struct SamplerComponent {  // Sampler.hh:128-149, 365-370
  std::string sampler_name;
  Index component_index = 0;
  bool operator<(SamplerComponent const &o) const {
    if (sampler_name == o.sampler_name) return component_index < o.component_index;
    return sampler_name < o.sampler_name;
  }
};

struct RequestedPrecision {  // Sampler.hh:151-180
  bool abs_convergence_is_required = false;
  double abs_precision = 0.0;
  bool rel_convergence_is_required = false;
  double rel_precision = 0.0;
};

// ---------------------------------------------------------------------------
// Statistics: math.hh:10-72, BasicStatistics.cc:24-131
// ---------------------------------------------------------------------------
inline double mean_of(std::vector<double> const &x) {
  // Eigen's mean() is sum()/size(); the sum is a reduction whose
  // association order is Eigen's (vectorised); here a sequential sum.  The
  // last-bit difference is why statistics are compared with a tolerance.
  double s = 0.0;
  for (double v : x) s += v;
  return s / static_cast<double>(x.size());
}

inline double variance(std::vector<double> const &x, double m) {  // math.hh:31-39
  double cov = 0.0;
  for (double v : x) {
    double d = v - m;
    cov += d * d;
  }
  return cov / static_cast<double>(x.size());
}

inline double approx_erf_inv(double x) {  // math.hh:62-72 (Winitzki, a = 0.147)
  const double one = 1.0;
  const double PI = 3.141592653589793238463;
  const double a = 0.147;
  double sgn = (x < 0.0) ? -one : one;
  double b = std::log((one - x) * (one + x));
  double c = 2.0 / (PI * a) + b * 0.5;
  double d = b / a;
  return sgn * std::sqrt(std::sqrt(c * c - d) - c);
}

// BasicStatistics.cc:24-48; `increment` is the interval between resampled
// observations (1.0 for unweighted ones)
inline double autocorrelation_factor(std::vector<double> const &x, double increment = 1.0) {
  Index N = x.size();
  double m = mean_of(x);
  double v0 = variance(x, m);
  if (std::abs(v0 / m) < 1e-8 || v0 == 0.0) return 1.0;
  for (Index i = 1; i < N; ++i) {
    Index range = N - i;
    double cov = 0.0;
    for (Index j = 0; j < range; ++j) cov += (x[j] - m) * (x[j + i] - m);
    cov /= range;
    if (std::abs(cov / v0) <= 0.5) {
      double rho = std::pow(2.0, (-1.0 / (i * increment)));
      return (1.0 + rho) / (1.0 - rho);
    }
  }
  return std::numeric_limits<double>::max();
}

// BasicStatistics.cc:50-73: the weighted series as a time series read at
// n equally spaced times i * W / n
inline std::vector<double> resample(std::vector<double> const &x, std::vector<double> const &w,
                                    double w_sum, Index n) {
  const double increment = w_sum / n;
  std::vector<double> out(n);
  Index j = 0;
  double W_j = 0.0;
  for (Index i = 0; i < n; ++i) {
    const double target = i * increment;
    while (W_j + w[j] < target) {
      W_j += w[j];
      ++j;
    }
    out[i] = x[j];
  }
  return out;
}

// misc/math.hh:45-54
inline double weighted_variance(std::vector<double> const &x, double m,
                                std::vector<double> const &w, double w_sum) {
  double cov = 0.0;
  for (size_t i = 0; i < x.size(); ++i) {
    const double d = x[i] - m;
    cov += w[i] * d * d;
  }
  return cov / w_sum;
}

struct BasicStatistics {  // BasicStatistics.hh:24-36
  double mean = 0.0;
  double calculated_precision = std::numeric_limits<double>::max();
};

inline BasicStatistics basic_statistics(std::vector<double> const &x,
                                        double confidence = 0.95) {  // BasicStatistics.cc:114-131
  if (x.empty()) throw std::runtime_error("observations.size()==0");
  BasicStatistics s;
  s.mean = mean_of(x);
  double v0 = variance(x, s.mean);
  double f = autocorrelation_factor(x);
  double fc = std::sqrt(2.0) * approx_erf_inv(confidence);
  s.calculated_precision = fc * std::sqrt(f * v0 / x.size());
  return s;
}

// BasicStatistics.cc:144-188 (weighted observations; method 1: weighted mean
// and variance, autocorrelation from the resampled series; method 2: all
// statistics from the resampled series).  Eigen's sum and dot are restated as
// sequential sums (their reduction order is Eigen's own).
struct WeightedStatistics {
  double mean = 0.0, calculated_precision = 0.0, variance = 0.0, autocorr = 0.0;
};
inline WeightedStatistics weighted_statistics(std::vector<double> const &x,
                                              std::vector<double> const &w, double confidence,
                                              int method, Index n_resamples) {
  if (x.empty()) throw std::runtime_error("observations.size()==0");
  if (x.size() != w.size()) throw std::runtime_error("observations.size() != sample_weight.size()");
  double W = 0.0;
  for (double v : w) W += v;
  const double increment = W / n_resamples;
  std::vector<double> eq = resample(x, w, W, n_resamples);
  const double fc = std::sqrt(2.0) * approx_erf_inv(confidence);
  WeightedStatistics s;
  if (method == 1) {
    double xw = 0.0;
    for (size_t i = 0; i < x.size(); ++i) xw += x[i] * w[i];
    s.mean = xw / W;
    s.variance = weighted_variance(x, s.mean, w, W);
    s.autocorr = autocorrelation_factor(eq, increment);
    s.calculated_precision = fc * std::sqrt(s.autocorr * s.variance / W);
  } else if (method == 2) {
    s.mean = mean_of(eq);
    s.variance = variance(eq, s.mean);
    s.autocorr = autocorrelation_factor(eq);
    s.calculated_precision = fc * std::sqrt(s.autocorr * s.variance / eq.size());
  } else {
    throw std::runtime_error("invalid method");
  }
  return s;
}

// EquilibrationCheck.cc:50-117
struct EquilibrationResult {
  bool is_equilibrated = false;
  CountType N_samples_for_equilibration = 0;
};

inline EquilibrationResult default_equilibration_check_impl(std::vector<double> const &obs,
                                                           double prec) {
  if (obs.empty()) throw std::runtime_error("observations.size()==0");
  EquilibrationResult r;
  double eps = (obs[0] == 0.0) ? 1e-8 : std::abs(obs[0]) * 1e-8;
  CountType N = obs.size();
  bool is_even = ((N % 2) == 0);
  bool all_same = true;
  for (CountType i = 0; i < N; i++)
    if (std::abs(obs[i] - obs[0]) > eps) {
      all_same = false;
      break;
    }
  if (all_same) {
    r.is_equilibrated = true;
    r.N_samples_for_equilibration = 0;
    return r;
  }
  CountType start1 = 0;
  CountType start2 = is_even ? N / 2 : (N / 2) + 1;
  double sum1 = 0.0, sum2 = 0.0;
  for (CountType i = 0; i < start2; ++i) sum1 += obs[i];
  for (CountType i = start2; i < N; ++i) sum2 += obs[i];
  while (std::abs((sum1 / (start2 - start1)) - (sum2 / (N - start2))) > prec &&
         start1 < N - 2) {
    if (is_even) {
      sum1 -= obs[start1];
      sum1 += obs[start2];
      sum2 -= obs[start2];
      start2++;
    } else {
      sum1 -= obs[start1];
    }
    start1++;
    is_even = !is_even;
  }
  double mean_tot = (sum1 + sum2) / (N - start1);
  if (obs[start1] < mean_tot) {
    while (obs[start1] < mean_tot && start1 < N - 1) start1++;
  } else {
    while (obs[start1] > mean_tot && start1 < N - 1) start1++;
  }
  r.is_equilibrated = (start1 < N - 1);
  r.N_samples_for_equilibration = start1;
  return r;
}

inline EquilibrationResult default_equilibration_check(std::vector<double> const &obs,
                                                      RequestedPrecision const &rp) {
  // EquilibrationCheck.cc:119-162 (unweighted branch)
  double prec;
  if (rp.abs_convergence_is_required) {
    prec = rp.abs_precision;
  } else if (rp.rel_convergence_is_required) {
    prec = std::abs(mean_of(obs) * rp.rel_precision);
  } else {
    EquilibrationResult r;
    r.is_equilibrated = true;
    return r;
  }
  return default_equilibration_check_impl(obs, prec);
}

struct EquilibrationCheckResults {  // EquilibrationCheck.hh:31-49
  bool all_equilibrated = false;
  CountType N_samples_for_all_to_equilibrate = 0;
  std::map<SamplerComponent, EquilibrationResult> individual_results;
};

inline EquilibrationCheckResults equilibration_check(
    std::map<SamplerComponent, RequestedPrecision> const &rp, SamplerMap const &samplers,
    bool check_all) {  // EquilibrationCheck.cc:180-225
  EquilibrationCheckResults res;
  if (!rp.size()) return res;
  res.all_equilibrated = true;
  for (auto const &p : rp) {
    auto it = samplers.find(p.first.sampler_name);
    if (it == samplers.end()) throw std::runtime_error("sampler not found");
    EquilibrationResult cur =
        default_equilibration_check(it->second->component(p.first.component_index), p.second);
    res.N_samples_for_all_to_equilibrate =
        std::max(res.N_samples_for_all_to_equilibrate, cur.N_samples_for_equilibration);
    res.all_equilibrated &= cur.is_equilibrated;
    res.individual_results.emplace(p.first, cur);
    if (!check_all && !res.all_equilibrated) break;
  }
  return res;
}

struct ConvergenceResult {  // ConvergenceCheck.hh:14-23
  bool is_converged = false;
  RequestedPrecision requested_precision;
  BasicStatistics stats;
};

struct ConvergenceCheckResults {  // ConvergenceCheck.hh:39-59
  bool all_converged = false;
  CountType N_samples_for_statistics = 0;
  std::map<SamplerComponent, ConvergenceResult> individual_results;
};

inline ConvergenceCheckResults convergence_check(
    SamplerMap const &samplers, std::map<SamplerComponent, RequestedPrecision> const &rp,
    CountType N_eq, double confidence) {  // ConvergenceCheck.hh:139-184
  ConvergenceCheckResults res;
  CountType N = get_n_samples(samplers);
  if (!rp.size()) {
    res.N_samples_for_statistics = N;
    return res;
  }
  if (N_eq >= N) return res;
  res.N_samples_for_statistics = N - N_eq;
  res.all_converged = true;
  for (auto const &p : rp) {
    auto it = samplers.find(p.first.sampler_name);
    if (it == samplers.end()) throw std::runtime_error("sampler not found");
    std::vector<double> comp = it->second->component(p.first.component_index);
    std::vector<double> tail(comp.end() - res.N_samples_for_statistics, comp.end());
    ConvergenceResult cur;
    cur.stats = basic_statistics(tail, confidence);
    cur.requested_precision = p.second;
    cur.is_converged = true;  // ConvergenceCheck.hh:74-90
    if (p.second.abs_convergence_is_required)
      cur.is_converged &= cur.stats.calculated_precision < p.second.abs_precision;
    if (p.second.rel_convergence_is_required)
      cur.is_converged &=
          std::abs(cur.stats.calculated_precision / cur.stats.mean) < p.second.rel_precision;
    res.all_converged &= cur.is_converged;
    res.individual_results.emplace(p.first, cur);
  }
  return res;
}

// CutoffCheck.hh:14-94
struct CutoffCheckParams {
  std::optional<CountType> min_count, min_sample, max_count, max_sample;
  std::optional<double> min_time, min_clocktime, max_time, max_clocktime;
};

inline bool all_minimums_met(CutoffCheckParams const &p, std::optional<CountType> count,
                             std::optional<double> time, CountType n_samples,
                             double clocktime) {
  if (p.min_sample.has_value() && n_samples < *p.min_sample) return false;
  if (p.min_count.has_value() && count.has_value() && *count < *p.min_count) return false;
  if (p.min_time.has_value() && time.has_value() && *time < *p.min_time) return false;
  if (p.min_clocktime.has_value() && clocktime < *p.min_clocktime) return false;
  return true;
}

inline bool any_maximum_met(CutoffCheckParams const &p, std::optional<CountType> count,
                            std::optional<double> time, CountType n_samples,
                            double clocktime) {
  if (p.max_sample.has_value() && n_samples >= *p.max_sample) return true;
  if (p.max_count.has_value() && count.has_value() && *count >= *p.max_count) return true;
  if (p.max_time.has_value() && time.has_value() && *time >= *p.max_time) return true;
  if (p.max_clocktime.has_value() && clocktime >= *p.max_clocktime) return true;
  return false;
}

struct CompletionCheckParams {  // CompletionCheck.hh:19-100
  CutoffCheckParams cutoff_params;
  std::map<SamplerComponent, RequestedPrecision> requested_precision;
  double confidence = 0.95;
  bool log_spacing = false;
  CountType check_begin = 100;
  CountType check_period = 100;
  double check_base = 10.0;
  double check_shift = 2.0;
  CountType check_period_max = 10000;

  CountType sample_check_linear(Index n) const { return check_begin + check_period * n; }
  CountType sample_check_log(Index n) const {
    return check_begin +
           static_cast<CountType>(std::round(std::pow(check_base, (n + check_shift))));
  }
  Index find_n_begin_linear() const {
    Index n = 0;
    auto delta = [&](Index k) { return sample_check_log(k) - sample_check_log(k - 1); };
    while (delta(n + 1) <= check_period_max) n += 1;
    return n;
  }
  CountType sample_check_log(Index n, Index n_begin_linear) const {
    if (n <= n_begin_linear) return sample_check_log(n);
    return sample_check_log(n_begin_linear) + check_period_max * (n - n_begin_linear);
  }
};

struct CompletionCheckResults {  // CompletionCheck.hh:103-171
  std::optional<CountType> count;
  double clocktime = 0.0;
  CountType n_samples = 0;
  bool has_all_minimums_met = false;
  bool has_any_maximum_met = false;
  std::optional<CountType> n_samples_at_convergence_check;
  EquilibrationCheckResults equilibration_check_results;
  ConvergenceCheckResults convergence_check_results;
  bool is_complete = false;
};

class CompletionCheck {  // CompletionCheck.hh:174-376
 public:
  explicit CompletionCheck(CompletionCheckParams p)
      : m_params(std::move(p)), m_n_begin_linear(m_params.find_n_begin_linear()) {
    m_t0 = std::chrono::steady_clock::now();
  }
  CompletionCheckParams const &params() const { return m_params; }
  CompletionCheckResults const &results() const { return m_results; }
  Index n_checks() const { return m_n_checks; }

  bool is_complete(SamplerMap const &samplers, std::optional<CountType> count) {
    CountType n_samples = get_n_samples(samplers);
    double clocktime = m_last_clocktime;
    if (n_samples != m_last_n_samples) {
      clocktime = time_s();
      m_last_n_samples = n_samples;
      m_last_clocktime = clocktime;
    }
    // partial_reset (CompletionCheck.hh:145-156)
    m_results.count = count;
    m_results.clocktime = clocktime;
    m_results.n_samples = n_samples;
    m_results.has_all_minimums_met = false;
    m_results.has_any_maximum_met = false;
    m_results.is_complete = false;

    m_results.has_all_minimums_met =
        all_minimums_met(m_params.cutoff_params, count, std::nullopt, n_samples, clocktime);
    if (!m_results.has_all_minimums_met) return false;
    m_results.has_any_maximum_met =
        any_maximum_met(m_params.cutoff_params, count, std::nullopt, n_samples, clocktime);
    if (m_results.has_any_maximum_met) {
      m_results.is_complete = true;
      if (n_samples != m_results.n_samples_at_convergence_check)
        check_convergence(samplers, n_samples);
      return true;
    }
    Index check_at = m_params.log_spacing
                         ? m_params.sample_check_log(m_n_checks, m_n_begin_linear)
                         : m_params.sample_check_linear(m_n_checks);
    if (n_samples >= check_at) {
      m_n_checks += 1;
      check_convergence(samplers, n_samples);
    }
    if (m_results.convergence_check_results.all_converged) m_results.is_complete = true;
    return m_results.is_complete;
  }

 private:
  void check_convergence(SamplerMap const &samplers, CountType n_samples) {
    if (m_params.requested_precision.size()) {
      m_results.n_samples_at_convergence_check = n_samples;
      m_results.equilibration_check_results =
          equilibration_check(m_params.requested_precision, samplers, false);
      if (m_results.equilibration_check_results.all_equilibrated) {
        m_results.convergence_check_results = convergence_check(
            samplers, m_params.requested_precision,
            m_results.equilibration_check_results.N_samples_for_all_to_equilibrate,
            m_params.confidence);
      } else {
        m_results.convergence_check_results = ConvergenceCheckResults();
      }
    }
  }
  double time_s() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - m_t0).count();
  }
  CompletionCheckParams m_params;
  CompletionCheckResults m_results;
  Index m_n_checks = 0;
  Index m_n_begin_linear = 0;
  Index m_last_n_samples = 0;
  double m_last_clocktime = 0.0;
  std::chrono::steady_clock::time_point m_t0;
};

// ---------------------------------------------------------------------------
// The main loop: basic_occupation_metropolis.hh:354-425 with the SGC Ising
// calculator of basic_semigrand_canonical.hh:415-470 and its three default
// sampling functions (:485-590).
// ---------------------------------------------------------------------------
struct TraceEntry {
  long site;
  int old_occ;
  int nbr_sum;
  int drew;       // 1 if a uniform was drawn (dE >= 0)
  int accepted;
};

struct SgcRunOutput {
  SamplerMap samplers;           // formation_energy, param_composition, potential_energy
  std::vector<long long> S, M;   // integer bond sum and magnetisation at each sample
  CountType n_pass = 0;
  BigCountType n_accept = 0, n_reject = 0;
  CompletionCheckResults results;
  Index n_checks = 0;
  std::vector<TraceEntry> trace;
  IsingConfiguration final_config;
};

inline long long bond_sum(IsingConfiguration const &c) {
  long long s = 0;
  for (Index j = 0; j < c.cols; ++j)
    for (Index i = 0; i < c.rows; ++i) {
      int v = c.occupation[i + c.rows * j];
      s += v * (c.occupation[c.within(i + 1, 0) + c.rows * j] +
                c.occupation[i + c.rows * c.within(j + 1, 1)]);
    }
  return s;
}

inline long long magnetisation(IsingConfiguration const &c) {
  long long m = 0;
  for (int v : c.occupation) m += v;
  return m;
}

inline SgcRunOutput sgc_basic_run(IsingConfiguration config, double J, bool use_nlist,
                                  double temperature, double mu,
                                  std::shared_ptr<std::mt19937_64> engine,
                                  CompletionCheckParams const &ccp, int sample_period,
                                  size_t trace_len = 0) {
  SgcRunOutput out;
  SemiGrandCanonicalPotential potential{IsingFormationEnergy(J, use_nlist)};
  potential.set_state(&config, mu);
  CountType n_steps_per_pass = config.n_sites;  // basic_semigrand_canonical.hh:432
  Index max_l = config.n_sites - 1;              // :304

  for (auto name : {"formation_energy", "param_composition", "potential_energy"})
    out.samplers.emplace(name, std::make_shared<Sampler>(1));
  CompletionCheck cc(ccp);

  double beta = 1.0 / (KB * temperature);  // basic_occupation_metropolis.hh:361
  RandomNumberGenerator rng(engine);
  Index n_pass_next_sample = sample_period;
  CountType n_step = 0;
  CountType n_pass = 0;
  while (!cc.is_complete(out.samplers, n_pass)) {  // :381-382
    long l = rng.random_int(max_l);                // propose, basic_semigrand_canonical.hh:308-315
    int new_occ = -config.occ(l);
    double dE = potential.delta(l, new_occ);  // :387
    bool drew = !(dE < 0.0);
    bool acc = metropolis_acceptance(dE, beta, rng);
    if (out.trace.size() < trace_len) {
      int nsum = 0;
      Index i = l % config.rows, j = l / config.rows;
      nsum = config.occ(config.to_linear(config.within(i + 1, 0), j)) +
             config.occ(config.to_linear(config.within(i - 1, 0), j)) +
             config.occ(config.to_linear(i, config.within(j + 1, 1))) +
             config.occ(config.to_linear(i, config.within(j - 1, 1)));
      out.trace.push_back({l, config.occ(l), nsum, drew ? 1 : 0, acc ? 1 : 0});
    }
    if (acc) {
      out.n_accept++;
      config.set_occ(l, new_occ);
    } else {
      out.n_reject++;
    }
    n_step++;
    if (n_step == n_steps_per_pass) {
      n_step = 0;
      n_pass += 1;
    }
    if (n_pass == n_pass_next_sample) {  // :406-421
      n_pass_next_sample += sample_period;
      out.samplers.at("formation_energy")->push_back({potential.formation_energy.per_unitcell()});
      out.samplers.at("param_composition")->push_back({potential.param_composition.per_unitcell()});
      out.samplers.at("potential_energy")->push_back({potential.per_unitcell()});
      out.S.push_back(bond_sum(config));
      out.M.push_back(magnetisation(config));
    }
  }
  out.n_pass = n_pass;
  out.results = cc.results();
  out.n_checks = cc.n_checks();
  out.final_config = config;
  return out;
}

// occupation_metropolis.hh:96-153 for the Ising SGC potential with the single
// site flip proposal: the FP64 running potential energy per unit cell
// (:106 initial value, :141 update per accepted step) after every pass.
inline std::vector<double> sgc_running_energy_run(IsingConfiguration config, double J,
                                                  bool use_nlist, double temperature, double mu,
                                                  std::shared_ptr<std::mt19937_64> engine,
                                                  long n_passes, double *initial = nullptr) {
  SemiGrandCanonicalPotential potential{IsingFormationEnergy(J, use_nlist)};
  potential.set_state(&config, mu);
  const double n_unitcells = static_cast<double>(config.n_sites);
  double pe = potential.per_supercell() / n_unitcells;
  if (initial) *initial = pe;
  const double beta = 1.0 / (KB * temperature);
  RandomNumberGenerator rng(engine);
  std::vector<double> out;
  for (long p = 0; p < n_passes; ++p) {
    for (long s = 0; s < config.n_sites; ++s) {
      long l = rng.random_int(config.n_sites - 1);
      int new_occ = -config.occ(l);
      double dE = potential.delta(l, new_occ);
      if (metropolis_acceptance(dE, beta, rng)) {
        config.set_occ(l, new_occ);
        pe += (dE / n_unitcells);
      }
    }
    out.push_back(pe);
  }
  return out;
}

// The loop of sgc_running_energy_run counted in steps, as a fixture sampling
// by step sees it (SamplingFixture.hh:154-169 with SAMPLE_MODE::BY_STEP,
// occupation_metropolis.hh:118-150): the running potential energy per unit
// cell and the magnetisation after each of n_steps steps (index 0: before the
// first step), and the final configuration.
struct StepRunOutput {
  std::vector<double> pe;
  std::vector<long long> M;
  IsingConfiguration final_config;
};
// after_step(g), when given, runs after step g (g = 1..n_steps) and may draw
// from the same engine (a fixture's stochastic sampling period does).
inline StepRunOutput sgc_step_run(IsingConfiguration config, double J, bool use_nlist,
                                  double temperature, double mu,
                                  std::shared_ptr<std::mt19937_64> engine, long n_steps,
                                  std::function<void(long)> after_step = nullptr) {
  SemiGrandCanonicalPotential potential{IsingFormationEnergy(J, use_nlist)};
  potential.set_state(&config, mu);
  const double n_unitcells = static_cast<double>(config.n_sites);
  double pe = potential.per_supercell() / n_unitcells;
  const double beta = 1.0 / (KB * temperature);
  RandomNumberGenerator rng(engine);
  StepRunOutput out;
  out.pe.push_back(pe);
  out.M.push_back(magnetisation(config));
  long long M = out.M.back();
  for (long s = 0; s < n_steps; ++s) {
    long l = rng.random_int(config.n_sites - 1);
    int new_occ = -config.occ(l);
    double dE = potential.delta(l, new_occ);
    if (metropolis_acceptance(dE, beta, rng)) {
      config.set_occ(l, new_occ);
      pe += (dE / n_unitcells);
      M += 2 * new_occ;
    }
    out.pe.push_back(pe);
    out.M.push_back(M);
    if (after_step) after_step(s + 1);
  }
  out.final_config = config;
  return out;
}

// ---------------------------------------------------------------------------
// Build-defined Ising models for the statistical modes: 2D square and 3D
// simple cubic lattices with nearest (J1) and next-nearest (J2) couplings,
// E_f = -J1 sum_NN s_i s_j - J2 sum_NNN s_i s_j, E_sgc = E_f - mu*(N+M)/2.
// Linear index: l = i + L0*(j + L1*k) (first index fastest, as in model.hh).
// ---------------------------------------------------------------------------
struct LatticeModel {
  int ndim = 2;
  long L[3] = {1, 1, 1};
  double J1 = 0.1, J2 = 0.0;
  long n_sites() const { return L[0] * L[1] * L[2]; }
  // NN offsets and NNN offsets (each bond listed once in the "forward" list)
  std::vector<std::array<int, 3>> nn_forward() const {
    if (ndim == 2) return {{1, 0, 0}, {0, 1, 0}};
    return {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  }
  std::vector<std::array<int, 3>> nnn_forward() const {
    if (ndim == 2) return {{1, 1, 0}, {1, -1, 0}};
    return {{1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
  }
  long wrap(long a, int d) const {
    long r = a % L[d];
    return r < 0 ? r + L[d] : r;
  }
  long index(long i, long j, long k) const {
    return wrap(i, 0) + L[0] * (wrap(j, 1) + L[1] * wrap(k, 2));
  }
  void coords(long l, long &i, long &j, long &k) const {
    i = l % L[0];
    j = (l / L[0]) % L[1];
    k = l / (L[0] * L[1]);
  }
};

// Neighbour sums at site l: nn and nnn (over all neighbours, both directions)
inline void neighbour_sums(LatticeModel const &m, std::vector<int8_t> const &s, long l,
                           int &nn, int &nnn) {
  long i, j, k;
  m.coords(l, i, j, k);
  nn = 0;
  nnn = 0;
  for (auto const &o : m.nn_forward()) {
    nn += s[m.index(i + o[0], j + o[1], k + o[2])];
    nn += s[m.index(i - o[0], j - o[1], k - o[2])];
  }
  if (m.J2 != 0.0) {
    for (auto const &o : m.nnn_forward()) {
      nnn += s[m.index(i + o[0], j + o[1], k + o[2])];
      nnn += s[m.index(i - o[0], j - o[1], k - o[2])];
    }
  }
}

// Integer observables: S1 = sum_NN s s, S2 = sum_NNN s s, M = sum s
inline void lattice_observables(LatticeModel const &m, std::vector<int8_t> const &s,
                                long long &S1, long long &S2, long long &M) {
  S1 = S2 = M = 0;
  long N = m.n_sites();
  for (long l = 0; l < N; ++l) {
    long i, j, k;
    m.coords(l, i, j, k);
    M += s[l];
    for (auto const &o : m.nn_forward()) S1 += s[l] * s[m.index(i + o[0], j + o[1], k + o[2])];
    for (auto const &o : m.nnn_forward()) S2 += s[l] * s[m.index(i + o[0], j + o[1], k + o[2])];
  }
}

// Sequential random-site Metropolis on a LatticeModel (the reference's
// single-site-flip SGC loop generalised to J2 and 3D), recording the integer
// observables after every pass.
inline void lattice_sequential_run(LatticeModel const &m, std::vector<int8_t> &s,
                                   double temperature, double mu, std::mt19937_64 &eng,
                                   long n_passes, std::vector<long long> &S1_out,
                                   std::vector<long long> &S2_out,
                                   std::vector<long long> &M_out) {
  RandomNumberGenerator rng(std::shared_ptr<std::mt19937_64>(&eng, [](std::mt19937_64 *) {}));
  double beta = 1.0 / (KB * temperature);
  long N = m.n_sites();
  long long S1, S2, M;
  lattice_observables(m, s, S1, S2, M);
  for (long p = 0; p < n_passes; ++p) {
    for (long step = 0; step < N; ++step) {
      long l = rng.random_int(N - 1);
      int nn, nnn;
      neighbour_sums(m, s, l, nn, nnn);
      int so = s[l];
      double ds = -2.0 * so;
      double dE = -m.J1 * ds * nn - m.J2 * ds * nnn - mu * (ds / 2.0);
      if (metropolis_acceptance(dE, beta, rng)) {
        s[l] = static_cast<int8_t>(-so);
        S1 += -2LL * so * nn;
        S2 += -2LL * so * nnn;
        M += -2LL * so;
      }
    }
    S1_out.push_back(S1);
    S2_out.push_back(S2);
    M_out.push_back(M);
  }
}

// ---------------------------------------------------------------------------
// Binary FCC cluster expansion (build-defined; the reference has no lattice
// model of its own beyond the 2D Ising one).
//
// Sites are the points (x, y, z) of a periodic 2L0 x 2L1 x 2L2 integer grid
// with x + y + z even (doubled conventional coordinates, L = conventional
// cells per edge, 4 L0 L1 L2 sites).  Linear index l = h + L0 * (y + 2 L1 z)
// with x = ((y + z) & 1) + 2h.  Occupation sigma in {0, 1}.
//
// Energy per supercell, correlations as orbit averages per primitive cell
// (one site) with per-site orbit multiplicities 6, 3, 12 (pairs of shells
// 1-3) and 8 (nearest-neighbour triangles):
//   E = N V0 + V1 N1 + (V2/6) P1 + (V3/3) P2 + (V4/12) P3 + (Vt/8) T
// with N1 = sum sigma, Pk = number of shell-k pairs with both sites occupied,
// T = number of nearest-neighbour triangles with all three occupied.
// The flip of site i (dsigma = 1 - 2 sigma_i) changes E by
//   dsigma * (V1 + (V2/6) n1 + (V3/3) n2 + (V4/12) n3 + (Vt/8) t)
// with nk = occupied shell-k neighbours and t = triangles through i whose two
// other sites are occupied.
// ---------------------------------------------------------------------------
struct FccModel {
  long L[3] = {3, 3, 3};
  double eci[6] = {0, 0, 0, 0, 0, 0};  // V0, V1, V2 (NN), V3 (2NN), V4 (3NN), Vt
  long n_sites() const { return 4 * L[0] * L[1] * L[2]; }

  // shell offsets by squared length in doubled coordinates: 2, 4, 6
  static std::vector<std::array<int, 3>> shell(int d2) {
    std::vector<std::array<int, 3>> out;
    for (int a = -2; a <= 2; ++a)
      for (int b = -2; b <= 2; ++b)
        for (int c = -2; c <= 2; ++c)
          if (((a + b + c) & 1) == 0 && a * a + b * b + c * c == d2) out.push_back({a, b, c});
    return out;
  }
  // unordered pairs of nearest-neighbour offsets that are nearest neighbours
  // of each other: the triangles through a site
  static std::vector<std::pair<int, int>> triangle_pairs() {
    auto nn = shell(2);
    std::vector<std::pair<int, int>> out;
    for (size_t a = 0; a < nn.size(); ++a)
      for (size_t b = a + 1; b < nn.size(); ++b) {
        int dx = nn[b][0] - nn[a][0], dy = nn[b][1] - nn[a][1], dz = nn[b][2] - nn[a][2];
        if (dx * dx + dy * dy + dz * dz == 2) out.push_back({(int)a, (int)b});
      }
    return out;
  }
  long wrap(long a, int d) const {
    long n = 2 * L[d], r = a % n;
    return r < 0 ? r + n : r;
  }
  long index(long x, long y, long z) const {
    x = wrap(x, 0);
    y = wrap(y, 1);
    z = wrap(z, 2);
    long h = (x - ((y + z) & 1)) / 2;
    return h + L[0] * (y + 2 * L[1] * z);
  }
  void coords(long l, long &x, long &y, long &z) const {
    long h = l % L[0], line = l / L[0];
    y = line % (2 * L[1]);
    z = line / (2 * L[1]);
    x = ((y + z) & 1) + 2 * h;
  }
};

struct FccCounts {
  long long N1 = 0, P1 = 0, P2 = 0, P3 = 0, T = 0;
};

inline FccCounts fcc_counts(FccModel const &m, std::vector<int8_t> const &s) {
  const auto s1 = FccModel::shell(2), s2 = FccModel::shell(4), s3 = FccModel::shell(6);
  const auto tp = FccModel::triangle_pairs();
  FccCounts c;
  long long p1 = 0, p2 = 0, p3 = 0, t = 0;
  const long N = m.n_sites();
  for (long l = 0; l < N; ++l) {
    if (!s[l]) continue;
    long x, y, z;
    m.coords(l, x, y, z);
    c.N1 += 1;
    auto occ = [&](std::array<int, 3> const &o) {
      return (long long)s[m.index(x + o[0], y + o[1], z + o[2])];
    };
    for (auto const &o : s1) p1 += occ(o);
    for (auto const &o : s2) p2 += occ(o);
    for (auto const &o : s3) p3 += occ(o);
    for (auto const &pr : tp) t += occ(s1[pr.first]) * occ(s1[pr.second]);
  }
  c.P1 = p1 / 2;  // every pair seen from both ends
  c.P2 = p2 / 2;
  c.P3 = p3 / 2;
  c.T = t / 3;  // every triangle seen from its three corners
  return c;
}

inline void fcc_local(FccModel const &m, std::vector<int8_t> const &s, long l, int &n1, int &n2,
                      int &n3, int &t) {
  static const auto s1 = FccModel::shell(2), s2 = FccModel::shell(4), s3 = FccModel::shell(6);
  static const auto tp = FccModel::triangle_pairs();
  long x, y, z;
  m.coords(l, x, y, z);
  auto occ = [&](std::array<int, 3> const &o) {
    return (int)s[m.index(x + o[0], y + o[1], z + o[2])];
  };
  n1 = n2 = n3 = t = 0;
  for (auto const &o : s1) n1 += occ(o);
  for (auto const &o : s2) n2 += occ(o);
  for (auto const &o : s3) n3 += occ(o);
  for (auto const &pr : tp) t += occ(s1[pr.first]) * occ(s1[pr.second]);
}

inline double fcc_energy(FccModel const &m, FccCounts const &c) {
  const double *v = m.eci;
  return v[0] * (double)m.n_sites() + v[1] * (double)c.N1 + v[2] / 6.0 * (double)c.P1 +
         v[3] / 3.0 * (double)c.P2 + v[4] / 12.0 * (double)c.P3 + v[5] / 8.0 * (double)c.T;
}

// Sequential random-site SGC Metropolis on the FCC model: the reference's
// loop (basic_occupation_metropolis.hh:381-403: propose a random site,
// dE of the SGC potential, metropolis_acceptance, apply) with the FCC
// energy, recording the integer counts after every pass.
inline void fcc_sequential_run(FccModel const &m, std::vector<int8_t> &s, double temperature,
                               double mu, std::mt19937_64 &eng, long n_passes,
                               std::vector<FccCounts> &out) {
  RandomNumberGenerator rng(std::shared_ptr<std::mt19937_64>(&eng, [](std::mt19937_64 *) {}));
  const double beta = 1.0 / (KB * temperature);
  const long N = m.n_sites();
  const double *v = m.eci;
  FccCounts c = fcc_counts(m, s);
  for (long p = 0; p < n_passes; ++p) {
    for (long step = 0; step < N; ++step) {
      long l = rng.random_int(N - 1);
      int n1, n2, n3, t;
      fcc_local(m, s, l, n1, n2, n3, t);
      const int ds = 1 - 2 * s[l];
      const double dE = ds * (v[1] + v[2] / 6.0 * n1 + v[3] / 3.0 * n2 + v[4] / 12.0 * n3 +
                              v[5] / 8.0 * t) -
                        mu * ds;
      if (metropolis_acceptance(dE, beta, rng)) {
        s[l] = static_cast<int8_t>(1 - s[l]);
        c.N1 += ds;
        c.P1 += ds * n1;
        c.P2 += ds * n2;
        c.P3 += ds * n3;
        c.T += ds * t;
      }
    }
    out.push_back(c);
  }
}

// ---------------------------------------------------------------------------
// Canonical (Kawasaki pair-swap) Metropolis: the generic occupation_metropolis
// loop (occupation_metropolis.hh:87-153) with propose_canonical_event
// (OccEventProposal.hh:109-236) over an OccLocation (OccLocation.cc:39-118,
// 216-283; OccLocation.hh:254-259), for a binary lattice with one asymmetric
// unit: occupant index o in {0, 1}, species index == occupant index,
// candidates [(0,0), (0,1)] (OccCandidate.cc:31-52) and the single canonical
// swap ((0,0),(0,1)) that make_canonical_swaps builds (OccCandidate.cc:99-110).
// The potential is the square-lattice Ising formation energy of model.hh
// evaluated on the spins s = 2o - 1, with the multi-site delta of
// model.hh:354-379 (each site's single delta in event order, the earlier
// sites applied).
// ---------------------------------------------------------------------------
struct OccMol {  // OccEvent.hh Mol: id, l, asym, species_index, loc
  Index id = 0, l = 0, asym = 0, species_index = 0, loc = 0;
};

struct BinaryOccLocation {
  std::vector<OccMol> mol;
  std::vector<Index> loc[2];  // m_loc[cand_index] -> mol ids
  std::vector<Index> l_to_mol;

  // OccLocation.cc:39-118 (every site mutates, so mol_id == l)
  void initialize(std::vector<int> const &occupation) {
    mol.clear();
    l_to_mol.clear();
    loc[0].clear();
    loc[1].clear();
    mol.resize(occupation.size());
    for (Index l = 0; l < (Index)occupation.size(); ++l) {
      Index cand = occupation[l];
      OccMol &m = mol[l];
      m.id = l;
      m.l = l;
      m.asym = 0;
      m.species_index = cand;
      m.loc = (Index)loc[cand].size();
      loc[cand].push_back(l);
      l_to_mol.push_back(l);
    }
  }
  Index cand_size(Index cand) const { return (Index)loc[cand].size(); }
  // OccLocation.hh:254-259: uniform over [0, size-1] with size_t bounds
  OccMol const &choose_mol(Index cand, std::mt19937_64 &engine) const {
    std::size_t maxv = loc[cand].size() - 1;
    std::size_t i = std::uniform_int_distribution<std::size_t>(0, maxv)(engine);
    return mol[loc[cand][i]];
  }
  // OccLocation.cc:254-283 (the occ_transform part; no atom tracking)
  void apply_transform(Index mol_id, Index from_species, Index to_species,
                       std::vector<int> &occupation) {
    OccMol &m = mol[mol_id];
    if (m.species_index != from_species)
      throw std::runtime_error("refcpu: species mismatch in apply");
    occupation[m.l] = (int)to_species;
    Index cand = m.species_index;
    Index back = loc[cand].back();
    loc[cand][m.loc] = back;
    mol[back].loc = m.loc;
    loc[cand].pop_back();
    m.species_index = to_species;
    cand = m.species_index;
    m.loc = (Index)loc[cand].size();
    loc[cand].push_back(m.id);
  }
};

struct CanonicalEvent {  // OccEvent: linear_site_index, new_occ, occ_transform
  Index mol_id[2], l[2], from[2], to[2];
  int new_occ[2];
};

// OccEventProposal.hh:109-152 (one swap type: the draw is made and the swap is
// always the first one), then :173-208
inline void propose_canonical_event(CanonicalEvent &e, BinaryOccLocation const &ol,
                                    std::mt19937_64 &engine) {
  double tsum0 = 0.0;
  double tsum1 = tsum0 + ((double)ol.cand_size(0)) * ((double)ol.cand_size(1));
  if (tsum1 == 0.0) throw std::runtime_error("Error in choose_canonical_swap: No events possible.");
  double rand = std::uniform_real_distribution<double>(0, tsum1)(engine);
  if (!(rand < tsum1)) throw std::runtime_error("Error in choose_canonical_swap");
  OccMol const &ma = ol.choose_mol(0, engine);
  e.mol_id[0] = ma.id;
  e.l[0] = ma.l;
  e.from[0] = 0;
  e.to[0] = 1;
  OccMol const &mb = ol.choose_mol(1, engine);
  e.mol_id[1] = mb.id;
  e.l[1] = mb.l;
  e.from[1] = 1;
  e.to[1] = 0;
  for (int i = 0; i < 2; ++i) e.new_occ[i] = (int)e.to[i];  // Conversions::occ_index
}

// model.hh:305-345 single-site delta on spins s = 2o - 1 (both variants give
// the same value: the four-neighbour sum is an exact integer)
inline double lattice_gas_single_delta(IsingConfiguration const &c, std::vector<int> const &occ,
                                       double J, Index l, int new_occ) {
  auto spin = [&](Index q) { return 2 * occ[q] - 1; };
  Index i = l % c.rows, j = l / c.rows;
  int sum = spin(c.to_linear(c.within(i + 1, 0), j)) + spin(c.to_linear(i, c.within(j + 1, 1))) +
            spin(c.to_linear(c.within(i - 1, 0), j)) + spin(c.to_linear(i, c.within(j - 1, 1)));
  return -J * ((2 * new_occ - 1) - spin(l)) * sum;
}

// model.hh:354-379
inline double lattice_gas_delta(IsingConfiguration const &c, std::vector<int> &occ, double J,
                                CanonicalEvent const &e) {
  double dE = 0.0;
  int orig[2];
  for (int i = 0; i < 2; ++i) {
    dE += lattice_gas_single_delta(c, occ, J, e.l[i], e.new_occ[i]);
    orig[i] = occ[e.l[i]];
    occ[e.l[i]] = e.new_occ[i];
  }
  for (int i = 0; i < 2; ++i) occ[e.l[i]] = orig[i];
  return dE;
}

struct CanonicalRunOutput {
  std::vector<double> e_run;      // potential energy per unit cell after each pass
  std::vector<long long> S;       // bond sum after each pass
  std::vector<int> final_occ;     // occupant indices
  BinaryOccLocation ol;
  BigCountType n_accept = 0, n_reject = 0;
  double initial = 0.0;
};

// occupation_metropolis.hh:96-153: n_passes passes of mol_size() steps
inline CanonicalRunOutput canonical_run(Index rows, Index cols, std::vector<int> occ, double J,
                                        double temperature, std::mt19937_64 &engine,
                                        long n_passes, BinaryOccLocation const *init_ol = nullptr,
                                        bool use_nlist = true) {
  CanonicalRunOutput out;
  IsingConfiguration shape(rows, cols, 1);
  IsingConfiguration spins(rows, cols, 1);
  for (Index l = 0; l < shape.n_sites; ++l) spins.occupation[l] = 2 * occ[l] - 1;
  IsingFormationEnergy fe(J, use_nlist);  // per_supercell variant (model.hh:259-290)
  fe.set_state(&spins);
  const double n_unitcells = static_cast<double>(shape.n_sites);
  double pe = fe.per_supercell() / n_unitcells;
  out.initial = pe;
  BinaryOccLocation ol;
  if (init_ol)
    ol = *init_ol;
  else
    ol.initialize(occ);
  const double beta = 1.0 / (KB * temperature);
  RandomNumberGenerator rng(std::shared_ptr<std::mt19937_64>(&engine, [](std::mt19937_64 *) {}));
  CanonicalEvent e;
  for (long p = 0; p < n_passes; ++p) {
    for (Index s = 0; s < (Index)ol.mol.size(); ++s) {
      propose_canonical_event(e, ol, engine);
      double dE = lattice_gas_delta(shape, occ, J, e);
      if (metropolis_acceptance(dE, beta, rng)) {
        out.n_accept++;
        for (int i = 0; i < 2; ++i) ol.apply_transform(e.mol_id[i], e.from[i], e.to[i], occ);
        pe += (dE / n_unitcells);
      } else {
        out.n_reject++;
      }
    }
    out.e_run.push_back(pe);
    for (Index l = 0; l < shape.n_sites; ++l) spins.occupation[l] = 2 * occ[l] - 1;
    out.S.push_back(bond_sum(spins));
  }
  out.final_occ = occ;
  out.ol = ol;
  return out;
}

// ---------------------------------------------------------------------------
// Occupant events over an OccLocation with several asymmetric units: the
// generic occupation_metropolis loop (occupation_metropolis.hh:87-153) with
// propose_semigrand_canonical_event (OccEventProposal.hh:261-352: one
// random_real over the summed candidate counts, then one choose_mol) or
// propose_canonical_event (:109-236: one random_real over the summed count
// products, then choose_mol on cand_a and on cand_b), OccLocation::initialize
// and apply (OccLocation.cc:39-118, 216-283; no atom tracking), the candidate
// list of OccCandidate.cc:31-52 and the swaps the caller gives (those of
// make_canonical_swaps / make_semigrand_canonical_swaps, OccCandidate.cc:
// 99-155, by default).  The system: a square lattice whose sites each belong
// to one asymmetric unit; each unit allows one or two occupants (species
// indices in occupant order); the configuration holds occupant indices.
// The potential is the square-lattice Ising formation energy on spins
// s = 2o - 1 of the occupant indices (model.hh:259-290, the multi-site delta
// of model.hh:354-379) minus the exchange potential of each species:
//   per_supercell: E_f, then - mu[species(l)] for l = 0 .. n-1 in site order;
//   delta, per transformed site k in event order, the earlier ones applied:
//     dE = (dE + (-J ((2 o' - 1) - s_l) sum_l)) - (mu[to] - mu[from]).
// ---------------------------------------------------------------------------
struct OccSystem {
  std::vector<int> site_asym;                  // Conversions::l_to_asym
  std::vector<std::vector<int>> asym_species;  // species index of each occupant index
  int n_species = 0;
  std::vector<std::pair<int, int>> cand;       // OccCandidateList: (asym, species)
  std::vector<std::vector<int>> cand_of;       // [asym][species] -> candidate index

  OccSystem(std::vector<int> _site_asym, std::vector<std::vector<int>> _asym_species,
            int _n_species)
      : site_asym(std::move(_site_asym)),
        asym_species(std::move(_asym_species)),
        n_species(_n_species) {
    // OccCandidate.cc:31-52: asymmetric units with more than one occupant, each
    // occupant in occupant order
    cand_of.assign(asym_species.size(), std::vector<int>(n_species, -1));
    for (int b = 0; b < (int)asym_species.size(); ++b) {
      if (asym_species[b].size() < 2) continue;
      for (int s : asym_species[b]) {
        cand_of[b][s] = (int)cand.size();
        cand.push_back({b, s});
      }
    }
  }
  int occ_size(int b) const { return (int)asym_species[b].size(); }
  int species_index(int b, int o) const { return asym_species[b][o]; }
  int occ_index(int b, int s) const {
    for (int o = 0; o < occ_size(b); ++o)
      if (asym_species[b][o] == s) return o;
    throw std::runtime_error("refcpu: species not allowed");
  }
  bool species_allowed(int b, int s) const {
    for (int x : asym_species[b])
      if (x == s) return true;
    return false;
  }
};

// OccCandidate.cc:99-110 (cand_a < cand_b in candidate order, species differ,
// each allowed on the other's unit) and :129-143 (same unit, species differ)
inline std::vector<std::pair<int, int>> occ_canonical_swaps(OccSystem const &sys) {
  std::vector<std::pair<int, int>> out;
  for (int a = 0; a < (int)sys.cand.size(); ++a)
    for (int b = 0; b < (int)sys.cand.size(); ++b) {
      auto ca = sys.cand[a], cb = sys.cand[b];
      if (ca < cb && ca.second != cb.second && sys.species_allowed(ca.first, cb.second) &&
          sys.species_allowed(cb.first, ca.second))
        out.push_back({a, b});
    }
  return out;
}
inline std::vector<std::pair<int, int>> occ_semigrand_swaps(OccSystem const &sys) {
  std::vector<std::pair<int, int>> out;
  for (int a = 0; a < (int)sys.cand.size(); ++a)
    for (int b = 0; b < (int)sys.cand.size(); ++b) {
      auto ca = sys.cand[a], cb = sys.cand[b];
      if (ca.first == cb.first && ca.second != cb.second && sys.species_allowed(ca.first, cb.second))
        out.push_back({a, b});
    }
  return out;
}

struct OccLocationG {
  OccSystem const *sys = nullptr;
  std::vector<OccMol> mol;
  std::vector<std::vector<Index>> loc;  // m_loc[cand_index] -> mol ids
  std::vector<Index> l_to_mol;

  // OccLocation.cc:39-118 without atoms: mutating sites in site order get mol
  // ids 0, 1, ...; fixed sites map to mol_size()
  void initialize(OccSystem const &s, std::vector<int> const &occupation) {
    sys = &s;
    mol.clear();
    l_to_mol.clear();
    loc.assign(s.cand.size(), {});
    Index nmut = 0;
    for (Index l = 0; l < (Index)occupation.size(); ++l)
      if (s.occ_size(s.site_asym[l]) > 1) ++nmut;
    mol.resize(nmut);
    Index id = 0;
    for (Index l = 0; l < (Index)occupation.size(); ++l) {
      const int b = s.site_asym[l];
      if (s.occ_size(b) > 1) {
        const int sp = s.species_index(b, occupation[l]);
        const int ci = s.cand_of[b][sp];
        OccMol &m = mol[id];
        m.id = id;
        m.l = l;
        m.asym = b;
        m.species_index = sp;
        m.loc = (Index)loc[ci].size();
        loc[ci].push_back(id);
        l_to_mol.push_back(id);
        ++id;
      } else {
        l_to_mol.push_back(nmut);
      }
    }
  }
  Index cand_size(int ci) const { return (Index)loc[ci].size(); }
  // OccLocation.hh:257-261
  OccMol const &choose_mol(int ci, std::mt19937_64 &engine) const {
    std::size_t maxv = loc[ci].size() - 1;
    std::size_t i = std::uniform_int_distribution<std::size_t>(0, maxv)(engine);
    return mol[loc[ci][i]];
  }
  // OccLocation.cc:254-283 (the occ_transform part)
  void apply_transform(Index mol_id, Index from_species, Index to_species,
                       std::vector<int> &occupation) {
    OccMol &m = mol[mol_id];
    if (m.species_index != from_species)
      throw std::runtime_error("refcpu: species mismatch in apply");
    occupation[m.l] = sys->occ_index((int)m.asym, (int)to_species);
    int ci = sys->cand_of[m.asym][m.species_index];
    Index back = loc[ci].back();
    loc[ci][m.loc] = back;
    mol[back].loc = m.loc;
    loc[ci].pop_back();
    m.species_index = to_species;
    ci = sys->cand_of[m.asym][m.species_index];
    m.loc = (Index)loc[ci].size();
    loc[ci].push_back(m.id);
  }
};

struct OccEventG {  // OccEvent: 1 (semi-grand) or 2 (canonical) transforms
  int n = 0;
  Index mol_id[2], l[2], from[2], to[2];
  int asym[2], new_occ[2];
};

// OccEventProposal.hh:109-152 / 261-303: the swap whose cumulative count range
// holds one random_real draw over the total
inline int occ_choose_swap(OccLocationG const &ol, std::vector<std::pair<int, int>> const &swaps,
                           bool canonical, std::mt19937_64 &engine) {
  std::vector<double> tsum(swaps.size() + 1);
  tsum[0] = 0.;
  for (size_t i = 0; i < swaps.size(); ++i)
    tsum[i + 1] = tsum[i] + (canonical ? ((double)ol.cand_size(swaps[i].first)) *
                                             ((double)ol.cand_size(swaps[i].second))
                                       : ((double)ol.cand_size(swaps[i].first)));
  if (tsum.back() == 0.0)
    throw std::runtime_error(canonical
                                 ? "Error in choose_canonical_swap: No events possible."
                                 : "Error in choose_semigrand_canonical_swap: No events possible.");
  double rand = std::uniform_real_distribution<double>(0, tsum.back())(engine);
  for (size_t i = 0; i < swaps.size(); ++i)
    if (rand < tsum[i + 1]) return (int)i;
  throw std::runtime_error("refcpu: no swap chosen");
}

// :173-208 (canonical, choose_mol on cand_a then cand_b) and :314-336
// (semi-grand, one choose_mol on cand_a)
inline void occ_propose(OccEventG &e, OccLocationG const &ol,
                        std::vector<std::pair<int, int>> const &swaps, bool canonical,
                        std::mt19937_64 &engine) {
  OccSystem const &s = *ol.sys;
  const auto &sw = swaps[occ_choose_swap(ol, swaps, canonical, engine)];
  const int ca[2] = {sw.first, sw.second}, cb[2] = {sw.second, sw.first};
  e.n = canonical ? 2 : 1;
  for (int i = 0; i < e.n; ++i) {
    OccMol const &m = ol.choose_mol(ca[i], engine);
    e.mol_id[i] = m.id;
    e.l[i] = m.l;
    e.asym[i] = s.cand[ca[i]].first;
    e.from[i] = s.cand[ca[i]].second;
    e.to[i] = s.cand[cb[i]].second;
  }
  for (int i = 0; i < e.n; ++i) e.new_occ[i] = s.occ_index(e.asym[i], (int)e.to[i]);
}

// the potential described above; occ is restored before returning
inline double occ_potential_delta(IsingConfiguration const &shape, std::vector<int> &occ, double J,
                                  std::vector<double> const &mu, OccEventG const &e) {
  double dE = 0.0;
  int orig[2];
  for (int i = 0; i < e.n; ++i) {
    dE = (dE + lattice_gas_single_delta(shape, occ, J, e.l[i], e.new_occ[i])) -
         (mu[e.to[i]] - mu[e.from[i]]);
    orig[i] = occ[e.l[i]];
    occ[e.l[i]] = e.new_occ[i];
  }
  for (int i = e.n - 1; i >= 0; --i) occ[e.l[i]] = orig[i];
  return dE;
}

inline double occ_potential_per_supercell(IsingConfiguration const &shape, OccSystem const &s,
                                          std::vector<int> const &occ, double J,
                                          std::vector<double> const &mu, bool use_nlist) {
  IsingConfiguration spins(shape.rows, shape.cols, 1);
  for (Index l = 0; l < shape.n_sites; ++l) spins.occupation[l] = 2 * occ[l] - 1;
  IsingFormationEnergy fe(J, use_nlist);
  fe.set_state(&spins);
  double E = fe.per_supercell();
  for (Index l = 0; l < shape.n_sites; ++l)
    E -= mu[s.species_index(s.site_asym[l], occ[l])];
  return E;
}

struct OccRunOutput {
  std::vector<double> e_run;   // potential energy per unit cell after each pass
  std::vector<long long> S;    // bond sum of the spins after each pass
  std::vector<long long> M;    // sum of the spins after each pass
  std::vector<int> final_occ;  // occupant indices
  OccLocationG ol;
  BigCountType n_accept = 0, n_reject = 0, n_steps = 0;
  double initial = 0.0;
  std::string error;  // what the proposal threw, if the loop stopped on it
};

// occupation_metropolis.hh:96-153 for n_steps steps (mol_size() steps per
// pass; a record after each completed pass)
inline OccRunOutput occ_run(Index rows, Index cols, OccSystem const &sys, std::vector<int> occ,
                            double J, std::vector<double> const &mu,
                            std::vector<std::pair<int, int>> const &swaps, bool canonical,
                            double temperature, double n_unitcells, std::mt19937_64 &engine,
                            long n_steps, OccLocationG const *init_ol = nullptr,
                            bool use_nlist = true) {
  OccRunOutput out;
  IsingConfiguration shape(rows, cols, 1);
  double pe = occ_potential_per_supercell(shape, sys, occ, J, mu, use_nlist) / n_unitcells;
  out.initial = pe;
  OccLocationG ol;
  if (init_ol) {
    ol = *init_ol;
    ol.sys = &sys;
  } else {
    ol.initialize(sys, occ);
  }
  const double beta = 1.0 / (KB * temperature);
  RandomNumberGenerator rng(std::shared_ptr<std::mt19937_64>(&engine, [](std::mt19937_64 *) {}));
  const long per_pass = (long)ol.mol.size();
  OccEventG e;
  IsingConfiguration spins(rows, cols, 1);
  for (long t = 0; t < n_steps; ++t) {
    try {
      occ_propose(e, ol, swaps, canonical, engine);
    } catch (std::exception const &x) {
      out.error = x.what();
      break;
    }
    double dE = occ_potential_delta(shape, occ, J, mu, e);
    if (metropolis_acceptance(dE, beta, rng)) {
      out.n_accept++;
      for (int i = 0; i < e.n; ++i) ol.apply_transform(e.mol_id[i], e.from[i], e.to[i], occ);
      pe += (dE / n_unitcells);
    } else {
      out.n_reject++;
    }
    out.n_steps++;
    if (per_pass > 0 && out.n_steps % per_pass == 0) {
      out.e_run.push_back(pe);
      long long m = 0;
      for (Index l = 0; l < shape.n_sites; ++l) {
        spins.occupation[l] = 2 * occ[l] - 1;
        m += spins.occupation[l];
      }
      out.S.push_back(bond_sum(spins));
      out.M.push_back(m);
    }
  }
  out.final_occ = occ;
  out.ol = ol;
  return out;
}

// ---------------------------------------------------------------------------
// Kinetic Monte Carlo with atom tracking: the loop of kinetic_monte_carlo.hh:
// 136-229 (select an event and its time increment, sample by time at every
// sample time the event's time reaches, apply the event through the
// OccLocation with atom tracking, advance the time, count the step) for a
// square lattice gas, with an event selector defined here (the reference
// takes its selector from outside the repository).
//
// Model: occupant 0 = vacancy, 1 = atom; E = -J * sum over nearest-neighbour
// bonds of n_i n_j; lattices of at least 3 x 3 sites.
// Selector (rejection kinetic Monte Carlo): attempts follow one another at
// 1/R, R = 4 * (number of atoms).  An attempt chooses an atom uniformly from
// the OccLocation list of atoms (choose_mol, OccLocation.hh:257-261) and a
// direction d = random_int(3) (0: +i, 1: -i, 2: +j, 3: -j); it is rejected if
// the target site holds an atom, otherwise dE = J (a - b) (a: atoms next to
// the origin other than at the target, b: atoms next to the target other
// than at the origin) and it is accepted without a draw when dE <= 0, else
// when random_real(1.0) < exp(-beta dE).  The accepted attempt is the selected
// event; its time increment is fl(k / R), k the attempts it took.
// The event (OccEvent.hh): occ_transform (origin: 1 -> 0), (target: 0 -> 1);
// atom_traj: the origin's atom to the target with delta_ijk the direction.
// Atoms are numbered in site order at the start (OccLocation.cc:68-106); each
// keeps its site, its displacement (the translation change since the start)
// and its jump count (OccLocation.cc:286-330).
// ---------------------------------------------------------------------------
struct KmcAtom {
  Index site = 0;
  long dx = 0, dy = 0, n_jumps = 0;
};

struct KmcRecord {  // the state a sample sees
  long events = 0;             // events applied
  double time = 0.0;           // sample time (by time) or the last event's time (by count)
  long long S = 0;             // sum over bonds of n_i n_j
  long long sum_sq = 0;        // sum over atoms of dx^2 + dy^2
  long long sum_dx = 0, sum_dy = 0;
};

struct KmcRunOutput {
  std::vector<KmcRecord> records;
  std::vector<int> final_occ;
  BinaryOccLocation ol;
  std::vector<KmcAtom> atoms;
  std::vector<Index> site_atom;  // atom at each site, -1 for a vacancy
  long events = 0;
  long long attempts = 0;
  double time = 0.0;
};

inline Index kmc_neighbour(Index l, int d, Index rows, Index cols) {
  const Index i = l % rows, j = l / rows;
  switch (d) {
    case 0: return (i + 1) % rows + rows * j;
    case 1: return (i + rows - 1) % rows + rows * j;
    case 2: return i + rows * ((j + 1) % cols);
    default: return i + rows * ((j + cols - 1) % cols);
  }
}

inline KmcRecord kmc_record(std::vector<int> const &occ, std::vector<KmcAtom> const &atoms,
                            Index rows, Index cols, long events, double time) {
  KmcRecord r;
  r.events = events;
  r.time = time;
  for (Index l = 0; l < rows * cols; ++l)
    if (occ[l]) r.S += occ[kmc_neighbour(l, 0, rows, cols)] + occ[kmc_neighbour(l, 2, rows, cols)];
  for (auto const &a : atoms) {
    r.sum_sq += a.dx * a.dx + a.dy * a.dy;
    r.sum_dx += a.dx;
    r.sum_dy += a.dy;
  }
  return r;
}

// kinetic_monte_carlo.hh:185-226 for n_events events.  time_period > 0:
// samples at the times k * time_period (k = 0, 1, ...; the sample index
// continues from next_k) taken before the event whose time reaches them;
// otherwise a sample after every `per_pass` events (by count, the state after
// the pass with the time of its last event).
inline KmcRunOutput kmc_run(Index rows, Index cols, std::vector<int> occ, double J,
                            double temperature, std::mt19937_64 &engine, long n_events,
                            double time_period = 0.0, long next_k = 0) {
  KmcRunOutput out;
  const Index n = rows * cols;
  out.ol.initialize(occ);
  out.site_atom.assign(n, -1);
  for (Index l = 0; l < n; ++l)
    if (occ[l]) {
      out.site_atom[l] = (Index)out.atoms.size();
      out.atoms.push_back(KmcAtom{l, 0, 0, 0});
    }
  const double beta = 1.0 / (KB * temperature);
  const double R = 4.0 * (double)out.atoms.size();
  const long per_pass = (long)n;
  RandomNumberGenerator rng(std::shared_ptr<std::mt19937_64>(&engine, [](std::mt19937_64 *) {}));
  double time = 0.0;
  long k_next = next_k;
  for (long ev = 0; ev < n_events; ++ev) {
    // select_event: attempts until one is accepted
    long k = 0;
    Index from = 0, to = 0;
    int dir = 0;
    while (true) {
      ++k;
      ++out.attempts;
      OccMol const &m = out.ol.choose_mol(1, engine);
      from = m.l;
      dir = (int)rng.random_int(3);
      to = kmc_neighbour(from, dir, rows, cols);
      if (occ[to]) continue;
      int a = 0, b = 0;
      for (int e = 0; e < 4; ++e) {
        const Index x = kmc_neighbour(from, e, rows, cols), y = kmc_neighbour(to, e, rows, cols);
        if (x != to) a += occ[x];
        if (y != from) b += occ[y];
      }
      const double dE = J * (double)(a - b);
      if (dE <= 0.0) break;
      if (rng.random_real(1.0) < std::exp(-dE * beta)) break;
    }
    const double event_time = time + (double)k / R;
    if (time_period > 0.0)
      while (event_time >= (double)k_next * time_period) {
        out.records.push_back(kmc_record(occ, out.atoms, rows, cols, ev, (double)k_next * time_period));
        ++k_next;
      }
    // OccLocation::apply with atom tracking: the two transforms in order,
    // then the atom's move
    out.ol.apply_transform(from, 1, 0, occ);
    out.ol.apply_transform(to, 0, 1, occ);
    const Index id = out.site_atom[from];
    out.site_atom[from] = -1;
    out.site_atom[to] = id;
    KmcAtom &at = out.atoms[id];
    at.site = to;
    at.dx += dir == 0 ? 1 : dir == 1 ? -1 : 0;
    at.dy += dir == 2 ? 1 : dir == 3 ? -1 : 0;
    at.n_jumps += 1;
    time = event_time;
    if (time_period <= 0.0 && (ev + 1) % per_pass == 0)
      out.records.push_back(kmc_record(occ, out.atoms, rows, cols, ev + 1, time));
  }
  out.final_occ = occ;
  out.events = n_events;
  out.time = time;
  return out;
}

// ---------------------------------------------------------------------------
// N-fold way (rejection-free) for the SGC Ising model: the loop of
// nfold.hh:78-149 (select an event and its time increment, sample by count
// before applying it with the increment as the sample weight, apply, advance
// time) with a class-based event selector defined here (the reference takes
// its selector from outside the repository).  Events are single-site flips;
// a site's rate is its Metropolis probability min(1, exp(-beta dE)) for the
// SGC potential delta of class k = (s > 0) * 5 + (nbr_sum + 4) / 2 (the
// exact mode's classes).  Sites are kept in one array partitioned by class
// (class k at [start[k], start[k+1])); per event:
//   R = sum_k n_k r_k (k = 0..9, FP64 in order);
//   class: the first k whose running sum exceeds u1 * R (u1 the canonical
//     double of one engine word; the last non-empty class if rounding leaves
//     none);
//   site: arr[start[k] + p], p uniform in [0, n_k) (one Lemire word);
//   dt = -log1p(-u3) / R (u3 the canonical double of one more word);
//   flip the site, then re-class the site and its neighbours (+i, -i, +j,
//   -j) in that order; a site moving from class a to b < a swaps with the
//   first entry of each segment it leaves, to b > a with the last.
// ---------------------------------------------------------------------------
struct NfoldChain {
  IsingConfiguration config;
  double J = 0.1, mu = 0.0, beta = 1.0;
  double rate[10];
  std::vector<long> arr, pos;
  long start[11];
  double time = 0.0;

  int class_of(long l) const {
    Index i = l % config.rows, j = l / config.rows;
    int nsum = config.occ(config.to_linear(config.within(i + 1, 0), j)) +
               config.occ(config.to_linear(config.within(i - 1, 0), j)) +
               config.occ(config.to_linear(i, config.within(j + 1, 1))) +
               config.occ(config.to_linear(i, config.within(j - 1, 1)));
    return (config.occ(l) > 0 ? 5 : 0) + (nsum + 4) / 2;
  }

  NfoldChain(IsingConfiguration c, double _J, double temperature, double _mu)
      : config(std::move(c)), J(_J), mu(_mu), beta(1.0 / (KB * temperature)) {
    for (int sidx = 0; sidx < 2; ++sidx)
      for (int m = 0; m < 5; ++m) {
        const int s_old = sidx ? 1 : -1, new_occ = -s_old, nsum = 2 * m - 4;
        // model.hh:312-314 and basic_semigrand_canonical.hh:188-191, as the exact mode
        const double dEf = -J * (new_occ - s_old) * nsum;
        const double Ndx = (new_occ - s_old) / 2.0;
        const double dE = dEf - mu * Ndx;
        rate[sidx * 5 + m] = dE < 0.0 ? 1.0 : std::exp(-dE * beta);
      }
    const long n = config.n_sites;
    arr.resize(n);
    pos.resize(n);
    long count[10] = {0};
    for (long l = 0; l < n; ++l) count[class_of(l)]++;
    start[0] = 0;
    for (int k = 0; k < 10; ++k) start[k + 1] = start[k] + count[k];
    long fill[10];
    for (int k = 0; k < 10; ++k) fill[k] = start[k];
    for (long l = 0; l < n; ++l) {
      const int k = class_of(l);
      arr[fill[k]] = l;
      pos[l] = fill[k]++;
    }
  }

  int class_at(long p) const {
    int k = 0;
    while (start[k + 1] <= p) ++k;
    return k;
  }

  void swap_at(long p, long q) {
    std::swap(arr[p], arr[q]);
    pos[arr[p]] = p;
    pos[arr[q]] = q;
  }

  void move(long l, int to) {
    int a = class_at(pos[l]);
    while (a < to) {  // become the first entry of segment a+1
      swap_at(pos[l], start[a + 1] - 1);
      start[a + 1]--;
      ++a;
    }
    while (a > to) {  // become the last entry of segment a-1
      swap_at(pos[l], start[a]);
      start[a]++;
      --a;
    }
  }

  // select an event: the site and its time increment (three words)
  double select(std::mt19937_64 &engine, long *site, double *total_rate = nullptr) const {
    double R = 0.0;
    for (int k = 0; k < 10; ++k) R += (double)(start[k + 1] - start[k]) * rate[k];
    if (total_rate) *total_rate = R;
    const double target = std::generate_canonical<double, 53>(engine) * R;
    double run = 0.0;
    int k = -1, last = -1;
    for (int q = 0; q < 10; ++q) {
      const long nq = start[q + 1] - start[q];
      if (nq == 0) continue;
      last = q;
      run += (double)nq * rate[q];
      if (target < run) {
        k = q;
        break;
      }
    }
    if (k < 0) k = last;
    const long nk = start[k + 1] - start[k];
    const long p = std::uniform_int_distribution<long>(0, nk - 1)(engine);
    *site = arr[start[k] + p];
    const double u3 = std::generate_canonical<double, 53>(engine);
    return -std::log1p(-u3) / R;
  }

  void apply(long l) {
    config.set_occ(l, -config.occ(l));
    Index i = l % config.rows, j = l / config.rows;
    const long nb[5] = {l, config.to_linear(config.within(i + 1, 0), j),
                        config.to_linear(config.within(i - 1, 0), j),
                        config.to_linear(i, config.within(j + 1, 1)),
                        config.to_linear(i, config.within(j - 1, 1))};
    for (long x : nb) move(x, class_of(x));
  }
};

struct NfoldRunOutput {
  std::vector<long long> S, M;  // state at each sample
  std::vector<double> time, weight;
  std::vector<double> rate;  // by pass: the total rate when the sample was taken
  std::vector<int> final_occ;
  std::vector<long> arr;
  double final_time = 0.0;
  long pending_site = -1;
  double pending_dt = 0.0;
};

// n_passes passes of N events (nfold.hh:112-146), sampled one of two ways:
//  * by pass (time_period <= 0): the sample of pass count p
//    (p % sample_period == 0) holds the state after p*N events and, as its
//    weight, the time increment of the event that follows (nfold.hh:93-107).
//    That event is selected at the end of pass p and applied first in pass
//    p+1, which leaves the stream order of one selection per event unchanged.
//    (Every event flips one spin, so with N even these samples only see
//    states of one flip parity; see tests/test_oracle_nfold.py.)
//  * by time (time_period > 0): samples at times k * time_period (k >= 1),
//    each holding the state in force at that time, weight 1
//    (nfold.hh:105-108, 128-129); the grid points inside one residence
//    interval are one record, weighted by their number, timed at the first.
//  * sample_start (by pass): also the sample at count 0, before any event,
//    weighted by the first event's increment (nfold.hh:120-122 when a
//    fixture's first sample count is 0).
inline NfoldRunOutput nfold_run(IsingConfiguration config, double J, double temperature,
                                double mu, std::mt19937_64 &engine, long n_passes,
                                long sample_period = 1, double time_period = 0.0,
                                bool sample_start = false) {
  NfoldChain ch(std::move(config), J, temperature, mu);
  NfoldRunOutput out;
  const long n = ch.config.n_sites;
  long pending = -1;
  double pending_dt = 0.0;
  long next_k = 1;
  auto sample_pass = [&]() {
    double R = 0.0;
    pending_dt = ch.select(engine, &pending, &R);
    out.S.push_back(bond_sum(ch.config));
    out.M.push_back(magnetisation(ch.config));
    out.time.push_back(ch.time);
    out.weight.push_back(pending_dt);
    out.rate.push_back(R);
  };
  if (time_period <= 0.0 && sample_start) sample_pass();
  for (long p = 1; p <= n_passes; ++p) {
    for (long s = 0; s < n; ++s) {
      long l;
      double dt;
      if (pending >= 0) {
        l = pending;
        dt = pending_dt;
        pending = -1;
      } else {
        dt = ch.select(engine, &l);
      }
      if (time_period > 0.0) {
        // first grid index k with k * period >= t_end (the predicate decides at the edge)
        const double t_end = ch.time + dt;
        long k = (long)std::ceil(t_end / time_period);
        if (k < next_k) k = next_k;
        while (k > next_k && (double)(k - 1) * time_period >= t_end) --k;
        while ((double)k * time_period < t_end) ++k;
        const long m = k - next_k;
        if (m > 0) {
          out.S.push_back(bond_sum(ch.config));
          out.M.push_back(magnetisation(ch.config));
          out.time.push_back((double)next_k * time_period);
          out.weight.push_back((double)m);
          next_k += m;
        }
      }
      ch.apply(l);
      ch.time += dt;
    }
    if (time_period <= 0.0 && p % sample_period == 0) sample_pass();
  }
  out.final_occ = ch.config.occupation;
  out.arr = ch.arr;
  out.final_time = ch.time;
  out.pending_site = pending;
  out.pending_dt = pending_dt;
  return out;
}

}  // namespace refcpu
