// Drop-in counterparts of libcasm-monte's completion checking:
//  CutoffCheckParams, all_minimums_met, any_maximum_met (checks/CutoffCheck.hh:14-94);
//  equilibration check (src/casm/monte/checks/EquilibrationCheck.cc:50-225);
//  convergence check (checks/ConvergenceCheck.hh:14-184);
//  CompletionCheckParams / Results / CompletionCheck (checks/CompletionCheck.hh:19-376).
// Same names, members, call order and decisions as the reference.
#ifndef CASM_MI355X_monte_CompletionCheck
#define CASM_MI355X_monte_CompletionCheck

#include "casm/monte/BasicStatistics.hh"
#include "casm/monte/MethodLog.hh"
#include "casm/monte/sampling/Sampler.hh"

namespace CASM {
namespace monte {

// --- cutoffs ------------------------------------------------------------------
struct CutoffCheckParams {
  std::optional<CountType> min_count;
  std::optional<TimeType> min_time;
  std::optional<CountType> min_sample;
  std::optional<TimeType> min_clocktime;
  std::optional<CountType> max_count;
  std::optional<TimeType> max_time;
  std::optional<CountType> max_sample;
  std::optional<TimeType> max_clocktime;
};

inline bool all_minimums_met(CutoffCheckParams const &p, std::optional<CountType> count,
                             std::optional<TimeType> time, CountType n_samples,
                             TimeType clocktime) {
  if (p.min_sample.has_value() && n_samples < p.min_sample.value()) return false;
  if (p.min_count.has_value() && count.has_value() && count.value() < p.min_count.value())
    return false;
  if (p.min_time.has_value() && time.has_value() && time.value() < p.min_time.value())
    return false;
  if (p.min_clocktime.has_value() && clocktime < p.min_clocktime.value()) return false;
  return true;
}

inline bool any_maximum_met(CutoffCheckParams const &p, std::optional<CountType> count,
                            std::optional<TimeType> time, CountType n_samples,
                            TimeType clocktime) {
  if (p.max_sample.has_value() && n_samples >= p.max_sample.value()) return true;
  if (p.max_count.has_value() && count.has_value() && count.value() >= p.max_count.value())
    return true;
  if (p.max_time.has_value() && time.has_value() && time.value() >= p.max_time.value())
    return true;
  if (p.max_clocktime.has_value() && clocktime >= p.max_clocktime.value()) return true;
  return false;
}

// --- equilibration ------------------------------------------------------------
struct IndividualEquilibrationCheckResult {
  bool is_equilibrated = false;
  CountType N_samples_for_equilibration = 0;
};

typedef std::function<IndividualEquilibrationCheckResult(
    VectorXd const &observations, VectorXd const &sample_weight,
    RequestedPrecision requested_precision)>
    EquilibrationCheckFunction;

/// Van de Walle & Asta two-partition test (EquilibrationCheck.cc:50-117).
inline IndividualEquilibrationCheckResult _default_equilibration_check(VectorXd const &obs,
                                                                       double prec) {
  if (obs.size() == 0)
    throw std::runtime_error("Error in equilibration_check: observations.size()==0");
  IndividualEquilibrationCheckResult result;
  double eps = (obs(0) == 0.0) ? 1e-8 : std::abs(obs(0)) * 1e-8;
  CountType N = obs.size();
  bool is_even = ((N % 2) == 0);
  bool all_same = true;
  for (CountType i = 0; i < N; i++)
    if (std::abs(obs(i) - obs(0)) > eps) {
      all_same = false;
      break;
    }
  if (all_same) {
    result.is_equilibrated = true;
    result.N_samples_for_equilibration = 0;
    return result;
  }
  CountType start1 = 0;
  CountType start2 = is_even ? N / 2 : (N / 2) + 1;
  double sum1 = obs.head(start2).sum();
  double sum2 = obs.segment(start2, N - start2).sum();
  while (std::abs((sum1 / (start2 - start1)) - (sum2 / (N - start2))) > prec &&
         start1 < N - 2) {
    if (is_even) {
      sum1 -= obs(start1);
      sum1 += obs(start2);
      sum2 -= obs(start2);
      start2++;
    } else {
      sum1 -= obs(start1);
    }
    start1++;
    is_even = !is_even;
  }
  double mean_tot = (sum1 + sum2) / (N - start1);
  if (obs(start1) < mean_tot) {
    while (obs(start1) < mean_tot && start1 < N - 1) start1++;
  } else {
    while (obs(start1) > mean_tot && start1 < N - 1) start1++;
  }
  result.is_equilibrated = (start1 < N - 1);
  result.N_samples_for_equilibration = start1;
  return result;
}

inline IndividualEquilibrationCheckResult default_equilibration_check(
    VectorXd const &obs, VectorXd const &w, RequestedPrecision rp) {
  double prec;
  if (rp.abs_convergence_is_required) {
    prec = rp.abs_precision;
  } else if (rp.rel_convergence_is_required) {
    prec = std::abs(obs.mean() * rp.rel_precision);
  } else {
    IndividualEquilibrationCheckResult r;
    r.is_equilibrated = true;
    return r;
  }
  if (w.size() == 0) return _default_equilibration_check(obs, prec);
  if (w.size() != obs.size())
    throw std::runtime_error(
        "Error in equilibration_check: sample_weight.size() != observations.size()");
  Index N = w.size();
  double W = 0.0;
  for (Index i = 0; i < N; ++i) W += w(i);
  double weight_factor = N / W;
  VectorXd wobs = obs;
  for (Index i = 0; i < N; ++i) wobs(i) *= weight_factor * w(i);
  return _default_equilibration_check(wobs, prec);
}

struct EquilibrationCheckResults {
  bool all_equilibrated = false;
  CountType N_samples_for_all_to_equilibrate = 0;
  std::map<SamplerComponent, IndividualEquilibrationCheckResult> individual_results;
};

inline EquilibrationCheckResults equilibration_check(
    EquilibrationCheckFunction f, std::map<SamplerComponent, RequestedPrecision> const &rp,
    SamplerMap const &samplers, Sampler const &sample_weight, bool check_all) {
  if (f == nullptr)
    throw std::runtime_error("Error in equilibration_check: equilibration_check_f == nullptr");
  EquilibrationCheckResults results;
  if (!rp.size()) return results;
  results.all_equilibrated = true;
  for (auto const &p : rp) {
    Sampler const &sampler = *find_or_throw(samplers, p.first)->second;
    IndividualEquilibrationCheckResult cur =
        f(sampler.component(p.first.component_index), sample_weight.component(0), p.second);
    results.N_samples_for_all_to_equilibrate =
        std::max(results.N_samples_for_all_to_equilibrate, cur.N_samples_for_equilibration);
    results.all_equilibrated &= cur.is_equilibrated;
    results.individual_results.emplace(p.first, cur);
    if (!check_all && !results.all_equilibrated) break;
  }
  return results;
}

// --- convergence --------------------------------------------------------------
template <typename StatisticsType>
struct IndividualConvergenceCheckResult {
  bool is_converged = false;
  RequestedPrecision requested_precision;
  StatisticsType stats;
};

template <typename StatisticsType>
IndividualConvergenceCheckResult<StatisticsType> convergence_check(
    StatisticsType const &stats, RequestedPrecision const &rp) {
  IndividualConvergenceCheckResult<StatisticsType> r;
  r.stats = stats;
  r.requested_precision = rp;
  r.is_converged = true;
  if (rp.abs_convergence_is_required)
    r.is_converged &= get_calculated_precision(stats) < rp.abs_precision;
  if (rp.rel_convergence_is_required)
    r.is_converged &= get_calculated_relative_precision(stats) < rp.rel_precision;
  return r;
}

template <typename StatisticsType>
struct ConvergenceCheckResults {
  bool all_converged = false;
  CountType N_samples_for_statistics = 0;
  std::map<SamplerComponent, IndividualConvergenceCheckResult<StatisticsType>>
      individual_results;
};

template <typename StatisticsType>
ConvergenceCheckResults<StatisticsType> convergence_check(
    SamplerMap const &samplers, Sampler const &sample_weight,
    std::map<SamplerComponent, RequestedPrecision> const &rp, CountType N_eq,
    CalcStatisticsFunction<StatisticsType> calc_statistics_f) {
  ConvergenceCheckResults<StatisticsType> results;
  CountType N = get_n_samples(samplers);
  if (!rp.size()) {
    results.N_samples_for_statistics = N;
    return results;
  }
  if (N_eq >= N) return results;
  results.N_samples_for_statistics = N - N_eq;
  results.all_converged = true;
  static const VectorXd empty_weight;
  for (auto const &p : rp) {
    Sampler const &sampler = *find_or_throw(samplers, p.first)->second;
    VectorXd tail = sampler.component(p.first.component_index)
                        .tail(results.N_samples_for_statistics);
    VectorXd wtail = sample_weight.n_samples() != 0
                         ? sample_weight.component(0).tail(results.N_samples_for_statistics)
                         : empty_weight;
    auto cur = convergence_check(calc_statistics_f(tail, wtail), p.second);
    results.all_converged &= cur.is_converged;
    results.individual_results.emplace(p.first, cur);
  }
  return results;
}

// --- completion ---------------------------------------------------------------
template <typename StatisticsType>
struct CompletionCheckParams {
  CompletionCheckParams()
      : equilibration_check_f(default_equilibration_check),
        calc_statistics_f(default_statistics_calculator<StatisticsType>()) {}
  CutoffCheckParams cutoff_params;
  EquilibrationCheckFunction equilibration_check_f;
  CalcStatisticsFunction<StatisticsType> calc_statistics_f;
  std::map<SamplerComponent, RequestedPrecision> requested_precision;
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

template <typename StatisticsType>
struct CompletionCheckResults {
  CompletionCheckParams<StatisticsType> params;
  std::optional<CountType> count;
  std::optional<TimeType> time;
  TimeType clocktime = 0.0;
  CountType n_samples = 0;
  bool has_all_minimums_met = false;
  bool has_any_maximum_met = false;
  std::optional<CountType> n_samples_at_convergence_check;
  EquilibrationCheckResults equilibration_check_results;
  ConvergenceCheckResults<StatisticsType> convergence_check_results;
  bool is_complete = false;

  void partial_reset(std::optional<CountType> _count = std::nullopt,
                     std::optional<TimeType> _time = std::nullopt, TimeType _clocktime = 0.0,
                     CountType _n_samples = 0) {
    count = _count;
    time = _time;
    clocktime = _clocktime;
    n_samples = _n_samples;
    has_all_minimums_met = false;
    has_any_maximum_met = false;
    is_complete = false;
  }
  void full_reset(std::optional<CountType> _count = std::nullopt,
                  std::optional<TimeType> _time = std::nullopt, TimeType _clocktime = 0.0,
                  CountType _n_samples = 0) {
    partial_reset(_count, _time, _clocktime, _n_samples);
    n_samples_at_convergence_check = std::nullopt;
    equilibration_check_results = EquilibrationCheckResults();
    convergence_check_results = ConvergenceCheckResults<StatisticsType>();
  }
};

template <typename StatisticsType>
class CompletionCheck {
 public:
  CompletionCheck(CompletionCheckParams<StatisticsType> params)
      : m_params(params), m_n_begin_linear(m_params.find_n_begin_linear()) {
    m_results.params = m_params;
    m_results.is_complete = false;
    if (m_params.equilibration_check_f == nullptr)
      throw std::runtime_error(
          "Error constructing CompletionCheck: params.equilibration_check_f == nullptr");
    if (m_params.calc_statistics_f == nullptr)
      throw std::runtime_error(
          "Error constructing CompletionCheck: params.calc_statistics_f == nullptr");
  }

  CompletionCheckParams<StatisticsType> const &params() const { return m_params; }
  CompletionCheckResults<StatisticsType> const &results() const { return m_results; }

  void reset() {
    m_results.full_reset();
    m_n_checks = 0;
    m_last_n_samples = 0;
    m_last_clocktime = 0.0;
  }

  bool is_complete(SamplerMap const &samplers, Sampler const &sample_weight, Log &log) {
    return _is_complete(samplers, sample_weight, std::nullopt, std::nullopt, log);
  }
  bool is_complete(SamplerMap const &samplers, Sampler const &sample_weight, CountType count,
                   Log &log) {
    return _is_complete(samplers, sample_weight, count, std::nullopt, log);
  }
  bool is_complete(SamplerMap const &samplers, Sampler const &sample_weight, TimeType time,
                   Log &log) {
    return _is_complete(samplers, sample_weight, std::nullopt, time, log);
  }
  bool is_complete(SamplerMap const &samplers, Sampler const &sample_weight, CountType count,
                   TimeType time, Log &log) {
    return _is_complete(samplers, sample_weight, count, time, log);
  }

  /// Number of periodic checks performed so far (m_n_checks).
  Index n_checks() const { return m_n_checks; }
  /// Sample count at which the next periodic check is due.
  CountType next_check_at() const {
    return m_params.log_spacing ? m_params.sample_check_log(m_n_checks, m_n_begin_linear)
                                : m_params.sample_check_linear(m_n_checks);
  }

 private:
  bool _is_complete(SamplerMap const &samplers, Sampler const &sample_weight,
                    std::optional<CountType> count, std::optional<TimeType> time, Log &log) {
    CountType n_samples = get_n_samples(samplers);
    TimeType clocktime = m_last_clocktime;
    if (n_samples != m_last_n_samples) {
      clocktime = log.time_s();
      m_last_n_samples = n_samples;
      m_last_clocktime = clocktime;
    }
    m_results.partial_reset(count, time, clocktime, n_samples);
    m_results.has_all_minimums_met =
        all_minimums_met(m_params.cutoff_params, count, time, n_samples, clocktime);
    if (!m_results.has_all_minimums_met) return false;
    m_results.has_any_maximum_met =
        any_maximum_met(m_params.cutoff_params, count, time, n_samples, clocktime);
    if (m_results.has_any_maximum_met) {
      m_results.is_complete = true;
      if (n_samples != m_results.n_samples_at_convergence_check)
        _check_convergence(samplers, sample_weight, n_samples);
      return true;
    }
    if (n_samples >= next_check_at()) {
      m_n_checks += 1;
      _check_convergence(samplers, sample_weight, n_samples);
    }
    if (m_results.convergence_check_results.all_converged) m_results.is_complete = true;
    return m_results.is_complete;
  }

  void _check_convergence(SamplerMap const &samplers, Sampler const &sample_weight,
                          CountType n_samples) {
    if (m_params.requested_precision.size()) {
      m_results.n_samples_at_convergence_check = n_samples;
      m_results.equilibration_check_results =
          equilibration_check(m_params.equilibration_check_f, m_params.requested_precision,
                              samplers, sample_weight, false);
      if (m_results.equilibration_check_results.all_equilibrated) {
        m_results.convergence_check_results = convergence_check<StatisticsType>(
            samplers, sample_weight, m_params.requested_precision,
            m_results.equilibration_check_results.N_samples_for_all_to_equilibrate,
            m_params.calc_statistics_f);
      } else {
        m_results.convergence_check_results = ConvergenceCheckResults<StatisticsType>();
      }
    }
  }

  CompletionCheckParams<StatisticsType> m_params;
  CompletionCheckResults<StatisticsType> m_results;
  Index m_n_checks = 0;
  Index m_n_begin_linear = 0;
  Index m_last_n_samples = 0;
  double m_last_clocktime = 0.0;
};

/// Fluent helper for requested precisions (RequestedPrecisionConstructor.hh:29,
/// python/libcasm/monte/sampling/_requested_precision_constructor.py:96):
/// converge(functions, params).set_precision("potential_energy", 0.001).
template <typename StatisticsType>
struct RequestedPrecisionConstructor {
  StateSamplingFunctionMap const &functions;
  CompletionCheckParams<StatisticsType> &params;
  RequestedPrecisionConstructor &set_precision(std::string const &name, double abs_value) {
    auto const &f = functions.at(name);
    for (Index i = 0; i < (Index)f.component_names.size(); ++i)
      params.requested_precision[SamplerComponent(name, i, f.component_names[i])] =
          RequestedPrecision::abs(abs_value);
    return *this;
  }
};

template <typename StatisticsType>
RequestedPrecisionConstructor<StatisticsType> converge(
    StateSamplingFunctionMap const &functions, CompletionCheckParams<StatisticsType> &params) {
  return RequestedPrecisionConstructor<StatisticsType>{functions, params};
}

}  // namespace monte
}  // namespace CASM

#endif
