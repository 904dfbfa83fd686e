// Drop-in counterparts of libcasm-monte's sampling containers:
//  Sampler (sampling/Sampler.hh:30-112, 225-326), SamplerComponent (:128-149),
//  RequestedPrecision (:151-180), get_n_samples (:410-416), find_or_throw
//  (:372-408), default component names (:333-363);
//  StateSamplingFunction (sampling/StateSamplingFunction.hh:20-54) and its
//  JSON-valued sibling (values kept as JSON text);
//  SamplingParams and sample_at (sampling/SamplingParams.hh:14-180);
//  stochastic sampling periods (SamplingParams.hh:186-272).
#ifndef CASM_MI355X_monte_Sampler
#define CASM_MI355X_monte_Sampler

#include "casm/monte/RandomNumberGenerator.hh"
#include "casm/monte/definitions.hh"

namespace CASM {
namespace monte {

inline std::vector<std::string> colmajor_component_names(Index n_rows, Index n_cols) {
  std::vector<std::string> r;
  for (Index i = 0; i < n_cols; ++i)
    for (Index j = 0; j < n_rows; ++j) r.push_back(std::to_string(j) + "," + std::to_string(i));
  return r;
}

inline std::vector<std::string> default_component_names(std::vector<Index> shape) {
  if (shape.size() == 0) return {"0"};
  if (shape.size() == 1) {
    std::vector<std::string> r;
    for (Index i = 0; i < shape[0]; ++i) r.push_back(std::to_string(i));
    return r;
  }
  if (shape.size() == 2) return colmajor_component_names(shape[0], shape[1]);
  throw std::runtime_error(
      "Error constructing sampler component names: >2 dimensions is not supported");
}

inline Index calc_n_components(std::vector<Index> const &shape) {
  Index r = 1;
  for (Index x : shape) r *= x;
  return r;
}

/// Vector-valued samples stored as one column per component (the reference
/// keeps a samples x components matrix and grows it by capacity_increment).
class Sampler {
 public:
  Sampler(std::vector<Index> shape, CountType capacity_increment = 1000)
      : Sampler(shape, default_component_names(shape), capacity_increment) {}
  Sampler(std::vector<Index> shape, std::vector<std::string> const &component_names,
          CountType capacity_increment = 1000)
      : m_n_components(component_names.size()),
        m_component_names(component_names),
        m_shape(shape),
        m_capacity_increment(capacity_increment) {
    clear();
  }

  void push_back(double const &value) {
    if (m_n_components != 1) throw std::runtime_error("Sampler::push_back: not a scalar sampler");
    grow_if_full();
    m_cols[0].push_back(value);
    ++m_n_samples;
  }
  void push_back(VectorXd const &v) {
    if (v.size() != m_n_components)
      throw std::runtime_error("Sampler::push_back: sample size != n_components");
    grow_if_full();
    for (Index c = 0; c < m_n_components; ++c) m_cols[c].push_back(v(c));
    ++m_n_samples;
  }
  void clear() {
    m_cols.assign(m_n_components, {});
    for (auto &c : m_cols) c.reserve(m_capacity_increment);
    m_capacity = m_capacity_increment;
    m_n_samples = 0;
  }
  void set_sample_capacity(CountType cap) {
    for (auto &c : m_cols) c.reserve(cap);
    m_capacity = cap;
  }
  void set_capacity_increment(CountType inc) { m_capacity_increment = inc; }
  std::vector<std::string> const &component_names() const { return m_component_names; }
  std::vector<Index> const &shape() const { return m_shape; }
  Index n_components() const { return m_n_components; }
  CountType n_samples() const { return m_n_samples; }
  CountType sample_capacity() const { return m_capacity; }
  /// All samples of one component (a column of the reference's values()).
  VectorXd component(Index i) const { return VectorXd(m_cols.at(i)); }
  /// One sample (a row of the reference's values()).
  VectorXd sample(CountType s) const {
    VectorXd r(m_n_components);
    for (Index c = 0; c < m_n_components; ++c) r(c) = m_cols[c].at(s);
    return r;
  }

 private:
  void grow_if_full() {
    if (m_n_samples == m_capacity) set_sample_capacity(m_capacity + m_capacity_increment);
  }
  Index m_n_components;
  std::vector<std::string> m_component_names;
  std::vector<Index> m_shape;
  CountType m_capacity_increment;
  CountType m_capacity = 0;
  CountType m_n_samples = 0;
  std::vector<std::vector<double>> m_cols;
};

typedef std::map<std::string, std::shared_ptr<Sampler>> SamplerMap;

struct jsonSampler {
  std::vector<std::string> values;  // JSON text per sample
};
typedef std::map<std::string, jsonSampler> jsonSamplerMap;

struct SamplerComponent {
  SamplerComponent(std::string _sampler_name, Index _component_index,
                   std::string _component_name)
      : sampler_name(_sampler_name),
        component_index(_component_index),
        component_name(_component_name) {}
  std::string sampler_name;
  Index component_index = 0;
  std::string component_name;
  bool operator<(SamplerComponent const &o) const {
    if (sampler_name == o.sampler_name) return component_index < o.component_index;
    return sampler_name < o.sampler_name;
  }
};

struct RequestedPrecision {
  bool abs_convergence_is_required = false;
  double abs_precision = 0.0;
  bool rel_convergence_is_required = false;
  double rel_precision = 0.0;
  static RequestedPrecision abs_and_rel(double a, double r) {
    RequestedPrecision x;
    x.abs_convergence_is_required = true;
    x.abs_precision = a;
    x.rel_convergence_is_required = true;
    x.rel_precision = r;
    return x;
  }
  static RequestedPrecision abs(double v) {
    RequestedPrecision x;
    x.abs_convergence_is_required = true;
    x.abs_precision = v;
    return x;
  }
  static RequestedPrecision rel(double v) {
    RequestedPrecision x;
    x.rel_convergence_is_required = true;
    x.rel_precision = v;
    return x;
  }
};

inline SamplerMap::const_iterator find_or_throw(SamplerMap const &samplers,
                                                SamplerComponent const &key) {
  auto it = samplers.find(key.sampler_name);
  if (it == samplers.end())
    throw std::runtime_error("Error finding sampler component: Sampler '" + key.sampler_name +
                             "' not found.");
  if (key.component_index >= it->second->n_components())
    throw std::runtime_error("Error finding sampler component: Requested component index " +
                             std::to_string(key.component_index) + ", but '" +
                             key.sampler_name + "' has " +
                             std::to_string(it->second->n_components()) + "components.");
  return it;
}

inline CountType get_n_samples(SamplerMap const &samplers) {
  if (samplers.size()) return samplers.begin()->second->n_samples();
  return 0;
}

/// A named function evaluated when a sample is taken.  `device_observable`
/// marks the functions whose values the GPU backend records itself (the
/// reference's default SGC sampling functions); any other function is called
/// on the host with the state mirror refreshed.
struct StateSamplingFunction {
  StateSamplingFunction(std::string _name, std::string _description, std::vector<Index> _shape,
                        std::function<VectorXd()> _function)
      : name(_name),
        description(_description),
        shape(_shape),
        component_names(default_component_names(_shape)),
        function(_function) {}
  StateSamplingFunction(std::string _name, std::string _description,
                        std::vector<std::string> const &_component_names,
                        std::vector<Index> _shape, std::function<VectorXd()> _function)
      : name(_name),
        description(_description),
        shape(_shape),
        component_names(_component_names),
        function(_function) {}
  std::string name, description;
  std::vector<Index> shape;
  std::vector<std::string> component_names;
  std::function<VectorXd()> function;
  int device_observable = -1;
  VectorXd operator()() const { return function(); }
};
typedef std::map<std::string, StateSamplingFunction> StateSamplingFunctionMap;

struct jsonStateSamplingFunction {
  jsonStateSamplingFunction(std::string _name, std::string _description,
                            std::function<std::string()> _function)
      : name(_name), description(_description), function(_function) {}
  std::string name, description;
  std::function<std::string()> function;
  std::string operator()() const { return function(); }
};
typedef std::map<std::string, jsonStateSamplingFunction> jsonStateSamplingFunctionMap;

struct SamplingParams {
  std::vector<std::string> sampler_names;
  std::vector<std::string> json_sampler_names;
  SAMPLE_MODE sample_mode = SAMPLE_MODE::BY_PASS;
  SAMPLE_METHOD sample_method = SAMPLE_METHOD::LINEAR;
  double period = 1.0;
  double begin = 1.0;
  double base = std::pow(10.0, 1.0 / 10.0);
  double shift = 10.0;
  std::function<double(CountType)> custom_sample_at;
  bool stochastic_sample_period = false;
  bool do_sample_trajectory = false;
  bool do_sample_time = false;
};

inline double sample_at(CountType sample_index, SamplingParams const &s) {
  double n = static_cast<double>(sample_index);
  if (s.sample_method == SAMPLE_METHOD::LINEAR) return s.begin + s.period * n;
  if (s.sample_method == SAMPLE_METHOD::LOG) return s.begin + std::pow(s.base, (n + s.shift));
  if (!s.custom_sample_at)
    throw std::runtime_error(
        "Error in sample_at: sample_method==SAMPLE_METHOD::CUSTOM and !custom_sample_at");
  return s.custom_sample_at(n);
}

/// Counts until the next sample of a stochastic sampling period
/// (SamplingParams.hh:186-204): the first dn >= 1 whose uniform draw on [0, 1)
/// falls below the rate, so dn is geometric with mean 1 / rate.
template <typename EngineType>
CountType stochastic_count_step(double sample_rate,
                                RandomNumberGenerator<EngineType> &random_number_generator) {
  CountType dn = 1;
  while (!(random_number_generator.random_real(1.0) < sample_rate)) ++dn;
  return dn;
}

/// Time until the next sample of a stochastic sampling period
/// (SamplingParams.hh:206-219): exponential with the given rate.
template <typename EngineType>
TimeType stochastic_time_step(TimeType sample_rate,
                              RandomNumberGenerator<EngineType> &random_number_generator) {
  return -std::log(random_number_generator.random_real(1.0)) / sample_rate;
}

/// Count or time of the sample_index-th sample with a stochastic period
/// (SamplingParams.hh:221-272): `begin` for the first, then the last sample's
/// count (time) plus a stochastic step at the rate 1 / period (LINEAR) or
/// 1 / (ln(base) base^(n + shift)) (LOG).  The reference computes no rate for
/// SAMPLE_METHOD::CUSTOM (its CUSTOM branch repeats the LOG test, so the rate
/// stays unset); that case is refused here.
template <typename EngineType>
double stochastic_sample_at(CountType sample_index, SamplingParams const &s,
                            RandomNumberGenerator<EngineType> &random_number_generator,
                            std::vector<CountType> const &sample_count,
                            std::vector<TimeType> const &sample_time) {
  if (sample_index == 0) return s.begin;
  const double n = static_cast<double>(sample_index);
  double rate;
  if (s.sample_method == SAMPLE_METHOD::LINEAR)
    rate = 1.0 / s.period;
  else if (s.sample_method == SAMPLE_METHOD::LOG)
    rate = 1.0 / (std::log(s.base) * std::pow(s.base, (n + s.shift)));
  else
    throw std::runtime_error(
        "Error in stochastic_sample_at: SAMPLE_METHOD::CUSTOM has no sampling rate");
  if (s.sample_mode == SAMPLE_MODE::BY_TIME)
    return sample_time.back() + stochastic_time_step(rate, random_number_generator);
  return sample_count.back() + stochastic_count_step(rate, random_number_generator);
}

}  // namespace monte
}  // namespace CASM

#endif
