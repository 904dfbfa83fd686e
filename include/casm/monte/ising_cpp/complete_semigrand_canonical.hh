// Drop-in counterpart of libcasm-monte's ising_cpp::complete_semigrand_canonical
// (include/casm/monte/ising_cpp/complete_semigrand_canonical.hh:30-839): the
// same SGC Ising calculator as basic_semigrand_canonical, with its own run-data
// structure (SemiGrandCanonicalData, :189-278) and the main loop written in
// `run` (:575-680) instead of in methods::basic_occupation_metropolis.  The
// loop's semantics are those of the basic calculator (one step per site and
// pass, n_steps_per_pass = n_sites, a completion check before every step,
// samples every sample_period passes, the status written when due after a
// sample and once at the end), so `run` drives the same device loop
// (ising_cpp/sgc_device_run.hh): occupations, sampled energies, counters,
// completion-check results and the engine are bit-identical to the
// reference's loop.
#ifndef CASM_MI355X_monte_ising_cpp_complete_semigrand_canonical
#define CASM_MI355X_monte_ising_cpp_complete_semigrand_canonical

#include <cmath>
#include <iostream>

#include "casm/casm_io/json/jsonParser.hh"
#include "casm/monte/checks/io/json/CompletionCheck_json_io.hh"
#include "casm/monte/ising_cpp/basic_semigrand_canonical.hh"
#include "casm/monte/ising_cpp/sgc_device_run.hh"

namespace CASM {
namespace monte {
namespace ising_cpp {
namespace complete_semigrand_canonical {

/// Conditions (temperature, exchange potential): :33-85.  The same class as the
/// basic calculator's, with its from_values/to_values.
using SemiGrandCanonicalConditions = basic_semigrand_canonical::SemiGrandCanonicalConditions;

/// E_sgc = E_f - mu . N x and its change for an event: :96-187.
template <typename SystemType>
using SemiGrandCanonicalPotential = basic_semigrand_canonical::SemiGrandCanonicalPotential<SystemType>;

/// Propose / apply single-site flips: :411-482.
template <typename EngineType>
using SemiGrandCanonicalEventGenerator =
    basic_semigrand_canonical::SemiGrandCanonicalEventGenerator<EngineType>;

/// Run data and results (:189-278): samplers, JSON samplers, the (empty)
/// sample weights, pass and step counts, accept/reject counts and the
/// completion check.
struct SemiGrandCanonicalData {
  typedef BasicStatistics statistics_type;

  SemiGrandCanonicalData(StateSamplingFunctionMap const &_sampling_functions,
                         jsonStateSamplingFunctionMap const &_json_sampling_functions,
                         CountType _n_steps_per_pass,
                         CompletionCheckParams<statistics_type> const &completion_check_params)
      : sampling_functions(_sampling_functions),
        json_sampling_functions(_json_sampling_functions),
        sample_weight({}),
        n_pass(0),
        n_steps_per_pass(_n_steps_per_pass),
        n_accept(0),
        n_reject(0),
        completion_check(completion_check_params) {
    for (auto const &p : sampling_functions)
      samplers.emplace(p.second.name,
                       std::make_shared<Sampler>(p.second.shape, p.second.component_names));
    for (auto const &p : json_sampling_functions) json_samplers.emplace(p.second.name, jsonSampler());
  }

  StateSamplingFunctionMap sampling_functions;
  SamplerMap samplers;
  jsonStateSamplingFunctionMap json_sampling_functions;
  jsonSamplerMap json_samplers;
  Sampler sample_weight;
  CountType n_pass;
  CountType n_steps_per_pass;
  BigCountType n_accept;
  BigCountType n_reject;
  CompletionCheck<statistics_type> completion_check;

  double acceptance_rate() const {
    const double a = static_cast<double>(n_accept), r = static_cast<double>(n_reject);
    return a / (a + r);
  }
  double rejection_rate() const {
    const double a = static_cast<double>(n_accept), r = static_cast<double>(n_reject);
    return r / (a + r);
  }
  /// Clear everything `run` sets (:266-277).
  void reset() {
    for (auto &p : samplers) p.second->clear();
    for (auto &p : json_samplers) p.second.values.clear();
    sample_weight.clear();
    n_pass = 0;
    n_accept = 0;
    n_reject = 0;
    completion_check.reset();
  }
};

/// :280-293
inline jsonParser &to_json(SemiGrandCanonicalData const &data, jsonParser &json) {
  json.put_obj();
  to_json(data.completion_check.results(), json["completion_check_results"]);
  json["n_pass"] = jsonParser(static_cast<long>(data.n_pass));
  json["n_steps_per_pass"] = jsonParser(static_cast<long>(data.n_steps_per_pass));
  json["n_accept"] = jsonParser(static_cast<long>(data.n_accept));
  json["n_reject"] = jsonParser(static_cast<long>(data.n_reject));
  json["acceptance_rate"] = jsonParser(data.acceptance_rate());
  json["rejection_rate"] = jsonParser(data.rejection_rate());
  return json;
}

/// Write nothing; start a new status lap (:299-304).
template <typename CalculatorType>
void write_no_status(CalculatorType const &, MethodLog &method_log) {
  method_log.log.begin_lap();
}

/// Status to stdout and the completion-check results to the method log
/// (:310-408): passes, samples, clock time and step rates; the current
/// parametric composition and formation energy; whether everything is
/// equilibrated and converged and, when not converged yet, each requested
/// precision against the calculated one.
template <typename CalculatorType>
void default_write_status(CalculatorType const &mc_calculator, MethodLog &method_log) {
  std::ostream &sout = std::cout;
  auto const &data = *mc_calculator.data;
  const double steps =
      static_cast<double>(data.n_pass) * static_cast<double>(data.n_steps_per_pass);
  const double time_s = method_log.log.time_s();
  sout << "Passes=" << data.n_pass << ", Samples=" << get_n_samples(data.samplers)
       << ", ClockTime(s)=" << time_s << ", Steps/Second=" << steps / time_s
       << ", Seconds/Step=" << time_s / steps << std::endl;
  sout << "  ParametricComposition="
       << mc_calculator.param_composition_calculator->per_unitcell()(0)
       << ", FormationEnergy=" << mc_calculator.formation_energy_calculator->per_unitcell()
       << std::endl;
  auto const &cc = data.completion_check;
  auto const &results = cc.results();
  auto finish = [&]() {
    method_log.reset();
    jsonParser json;
    to_json(results, json);
    std::ostringstream os;
    json.print(os);
    method_log.log << os.str() << std::endl;
    method_log.log.begin_lap();
  };
  const bool all_equilibrated = results.equilibration_check_results.all_equilibrated;
  sout << "  AllEquilibrated=" << all_equilibrated << std::endl;
  if (!all_equilibrated) return finish();
  const bool all_converged = results.convergence_check_results.all_converged;
  sout << "  AllConverged=" << all_converged << std::endl;
  if (all_converged) return finish();
  auto const &individual = results.convergence_check_results.individual_results;
  for (auto const &p : cc.params().requested_precision) {
    auto it = individual.find(p.first);
    if (it == individual.end()) continue;
    const double abs_prec = it->second.stats.calculated_precision;
    const double mean = it->second.stats.mean;
    const double rel_prec = std::fabs(abs_prec / mean);
    auto const &req = p.second;
    if (req.abs_convergence_is_required)
      sout << "  - " << p.first.sampler_name << "(" << p.first.component_index
           << "): mean=" << mean << ", abs_prec=" << abs_prec
           << " < requested=" << req.abs_precision << " == "
           << bool(abs_prec < req.abs_precision) << std::endl;
    if (req.rel_convergence_is_required)
      sout << "  - " << p.first.sampler_name << "(" << p.first.component_index
           << "): mean=" << mean << ", rel_prec=" << rel_prec
           << " < requested=" << req.rel_precision << " == "
           << bool(rel_prec < req.rel_precision) << std::endl;
  }
  finish();
}

/// The calculator (:485-681).  `run` returns nothing; results stay in `data`.
template <typename SystemType, typename EventGeneratorType>
class SemiGrandCanonicalCalculator {
 public:
  typedef SystemType system_type;
  typedef typename system_type::state_type state_type;
  typedef typename system_type::formation_energy_f_type formation_energy_f_type;
  typedef typename system_type::param_composition_f_type param_composition_f_type;
  typedef SemiGrandCanonicalPotential<system_type> potential_type;
  typedef EventGeneratorType event_generator_type;
  typedef typename event_generator_type::engine_type engine_type;
  typedef typename event_generator_type::random_number_generator_type random_number_generator_type;

  SemiGrandCanonicalCalculator(std::shared_ptr<system_type> _system)
      : system(throw_if_null(_system,
                             "Error constructing SemiGrandCanonicalCalculator: _system==nullptr")),
        state(nullptr),
        conditions(nullptr),
        potential(_system),
        formation_energy_calculator(&potential.formation_energy_calculator),
        param_composition_calculator(&potential.param_composition_calculator) {}

  std::shared_ptr<system_type> system;
  state_type *state;
  std::shared_ptr<SemiGrandCanonicalConditions> conditions;
  potential_type potential;
  formation_energy_f_type *formation_energy_calculator;
  param_composition_f_type *param_composition_calculator;
  std::shared_ptr<SemiGrandCanonicalData> data;

  /// The device record of the sample being taken (set while sampling functions
  /// are evaluated during run).
  std::optional<mcg_sample> current_record;

  /// Observable of the current sample from the device's exact integers, with
  /// the reference's FP64 rounding.
  double device_observable(int which) const {
    return detail::record_observable(*this, *current_record, which);
  }

  template <typename WriteStatusF = void (*)(SemiGrandCanonicalCalculator const &, MethodLog &)>
  void run(state_type &_state, StateSamplingFunctionMap const &sampling_functions,
           jsonStateSamplingFunctionMap const &json_sampling_functions,
           CompletionCheckParams<BasicStatistics> const &completion_check_params,
           event_generator_type event_generator, int sample_period = 1,
           std::optional<MethodLog> method_log = std::nullopt,
           std::shared_ptr<engine_type> random_engine = nullptr,
           WriteStatusF write_status_f = default_write_status<SemiGrandCanonicalCalculator>) {
    // set-up (:575-606)
    state = &_state;
    conditions = std::make_shared<SemiGrandCanonicalConditions>(
        SemiGrandCanonicalConditions::from_values(state->conditions));
    const CountType n_steps_per_pass = state->configuration.n_sites;
    potential.set_state(state, conditions);
    event_generator.set_state(state);
    data = std::make_shared<SemiGrandCanonicalData>(sampling_functions, json_sampling_functions,
                                                    n_steps_per_pass, completion_check_params);
    // the main loop (:617-678) on the device
    detail::run_sgc_on_device(*this, *data, sample_period, method_log, random_engine,
                              write_status_f);
  }
};

/// Sampling functions "param_composition", "formation_energy",
/// "potential_energy" and the JSON "configuration" (:694-839): the same
/// functions as the basic calculator's (they read the calculator's current
/// device record during run, its calculators otherwise).
using basic_semigrand_canonical::make_configuration_json_f;
using basic_semigrand_canonical::make_formation_energy_f;
using basic_semigrand_canonical::make_parametric_composition_f;
using basic_semigrand_canonical::make_potential_energy_f;

}  // namespace complete_semigrand_canonical
}  // namespace ising_cpp
}  // namespace monte
}  // namespace CASM

#endif
