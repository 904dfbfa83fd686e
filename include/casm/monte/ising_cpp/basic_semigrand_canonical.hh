// Drop-in counterpart of libcasm-monte's SGC Ising calculator
// (ising_cpp/basic_semigrand_canonical.hh:36-627) and the data/loop of
// methods/basic_occupation_metropolis.hh:18-425, with the main loop on the
// MI355X (the loop and the argument for running it in device launches are in
// ising_cpp/sgc_device_run.hh).
//
// SemiGrandCanonicalCalculator::run keeps the reference signature and
// semantics: the same engine stream is consumed (the shared engine is loaded
// into the device chain and written back at the end), the same samples are
// taken, and the completion check is called with the same arguments at every
// point where its answer can change.
#ifndef CASM_MI355X_monte_ising_cpp_basic_semigrand_canonical
#define CASM_MI355X_monte_ising_cpp_basic_semigrand_canonical

#include "casm/monte/RandomNumberGenerator.hh"
#include "casm/monte/checks/CompletionCheck.hh"
#include "casm/monte/events/OccEvent.hh"
#include "casm/monte/gpu/Handle.hh"
#include "casm/monte/ising_cpp/model.hh"
#include "casm/monte/ising_cpp/sgc_device_run.hh"

namespace CASM {
namespace monte {

namespace methods {

/// basic_occupation_metropolis.hh:18-115
template <typename StatisticsType>
struct BasicOccupationMetropolisData {
  typedef StatisticsType statistics_type;
  BasicOccupationMetropolisData(StateSamplingFunctionMap const &_sampling_functions,
                                jsonStateSamplingFunctionMap const &_json_sampling_functions,
                                CountType _n_steps_per_pass,
                                CompletionCheckParams<statistics_type> const &ccp)
      : sampling_functions(_sampling_functions),
        json_sampling_functions(_json_sampling_functions),
        sample_weight({}),
        n_steps_per_pass(_n_steps_per_pass),
        completion_check(ccp) {
    for (auto const &p : sampling_functions)
      samplers.emplace(p.second.name,
                       std::make_shared<Sampler>(p.second.shape, p.second.component_names));
    for (auto const &p : json_sampling_functions) json_samplers.emplace(p.second.name, jsonSampler());
    n_pass = 0;
    n_accept = 0;
    n_reject = 0;
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
    double a = static_cast<double>(n_accept), r = static_cast<double>(n_reject);
    return a / (a + r);
  }
  double rejection_rate() const {
    double a = static_cast<double>(n_accept), r = static_cast<double>(n_reject);
    return r / (a + r);
  }
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

/// Passes, samples, clock time and rates (basic_occupation_metropolis.hh:223-241).
template <typename StatisticsType>
void default_write_run_status(BasicOccupationMetropolisData<StatisticsType> const &data,
                              MethodLog &method_log, std::ostream &sout) {
  double steps = static_cast<double>(data.n_pass) * static_cast<double>(data.n_steps_per_pass);
  double time_s = method_log.log.time_s();
  sout << "Passes=" << data.n_pass << ", Samples=" << get_n_samples(data.samplers)
       << ", ClockTime(s)=" << time_s << ", Steps/Second=" << steps / time_s
       << ", Seconds/Step=" << time_s / steps << std::endl;
}

}  // namespace methods

namespace ising_cpp {
namespace basic_semigrand_canonical {

class SemiGrandCanonicalConditions {
 public:
  SemiGrandCanonicalConditions() {}
  SemiGrandCanonicalConditions(double _temperature, VectorXd _exchange_potential)
      : temperature(_temperature), exchange_potential(_exchange_potential) {}
  double temperature = 0.0;
  VectorXd exchange_potential;
  static SemiGrandCanonicalConditions from_values(ValueMap const &values) {
    if (!values.scalar_values.count("temperature"))
      throw std::runtime_error("Missing required condition: \"temperature\"");
    if (!values.vector_values.count("exchange_potential"))
      throw std::runtime_error("Missing required condition: \"exchange_potential\"");
    return SemiGrandCanonicalConditions(values.scalar_values.at("temperature"),
                                        values.vector_values.at("exchange_potential"));
  }
  ValueMap to_values() const {
    ValueMap v;
    v.scalar_values["temperature"] = temperature;
    v.vector_values["exchange_potential"] = exchange_potential;
    return v;
  }
};

template <typename SystemType>
class SemiGrandCanonicalPotential {
 public:
  typedef SystemType system_type;
  typedef typename system_type::state_type state_type;
  typedef typename system_type::formation_energy_f_type formation_energy_f_type;
  typedef typename system_type::param_composition_f_type param_composition_f_type;

  SemiGrandCanonicalPotential(std::shared_ptr<system_type> _system)
      : system(throw_if_null(_system,
                             "Error constructing SemiGrandCanonicalPotential: _system==nullptr")),
        state(nullptr),
        conditions(nullptr),
        formation_energy_calculator(system->formation_energy_calculator),
        param_composition_calculator(system->param_composition_calculator) {}
  std::shared_ptr<system_type> system;
  state_type const *state;
  std::shared_ptr<SemiGrandCanonicalConditions> conditions;
  formation_energy_f_type formation_energy_calculator;
  param_composition_f_type param_composition_calculator;

  void set_state(state_type const *_state,
                 std::shared_ptr<SemiGrandCanonicalConditions> _conditions) {
    state = throw_if_null(_state, "Error in SemiGrandCanonicalPotential::set_state: _state is nullptr");
    conditions = throw_if_null(
        _conditions, "Error in SemiGrandCanonicalPotential::set_state: _conditions is nullptr");
    formation_energy_calculator.set_state(_state);
    param_composition_calculator.set_state(_state);
  }
  double per_supercell() {
    return formation_energy_calculator.per_supercell() -
           conditions->exchange_potential.dot(param_composition_calculator.per_supercell());
  }
  double per_unitcell() { return per_supercell() / state->configuration.n_unitcells; }
  double occ_delta_per_supercell(std::vector<Index> const &lsi,
                                 std::vector<int> const &new_occ) const {
    double dEf = formation_energy_calculator.occ_delta_per_supercell(lsi, new_occ);
    VectorXd Ndx = param_composition_calculator.occ_delta_per_supercell(lsi, new_occ);
    return dEf - conditions->exchange_potential.dot(Ndx);
  }
  double occ_delta_per_supercell(OccEvent const &e) const {
    return occ_delta_per_supercell(e.linear_site_index, e.new_occ);
  }
};

struct SemiGrandCanonicalData : public methods::BasicOccupationMetropolisData<BasicStatistics> {
  SemiGrandCanonicalData(StateSamplingFunctionMap const &f, jsonStateSamplingFunctionMap const &jf,
                         CountType n_steps_per_pass,
                         CompletionCheckParams<BasicStatistics> const &ccp)
      : methods::BasicOccupationMetropolisData<BasicStatistics>(f, jf, n_steps_per_pass, ccp) {}
};

/// Single-site flip events (basic_semigrand_canonical.hh:267-321).  The GPU
/// loop proposes on the device with the same draw; this class keeps the
/// reference interface for host-side use.
template <typename EngineType>
class SemiGrandCanonicalEventGenerator {
 public:
  typedef IsingState state_type;
  typedef EngineType engine_type;
  typedef RandomNumberGenerator<engine_type> random_number_generator_type;
  SemiGrandCanonicalEventGenerator() : state(nullptr), m_max_linear_site_index(0) {
    occ_event.linear_site_index = {0};
    occ_event.new_occ = {1};
  }
  state_type *state;
  OccEvent occ_event;
  void set_state(state_type *_state) {
    state = throw_if_null(_state, "Error in SemiGrandCanonicalEventGenerator::set_state: _state==nullptr");
    m_max_linear_site_index = state->configuration.n_sites - 1;
  }
  OccEvent const &propose(random_number_generator_type &rng) {
    occ_event.linear_site_index[0] = rng.random_int(m_max_linear_site_index);
    occ_event.new_occ[0] = -state->configuration.occ(occ_event.linear_site_index[0]);
    return occ_event;
  }
  void apply(OccEvent const &e) { state->configuration.set_occ(e.linear_site_index[0], e.new_occ[0]); }

 private:
  Index m_max_linear_site_index;
};

template <typename CalculatorType>
void write_no_status(CalculatorType const &, MethodLog &method_log) {
  method_log.log.begin_lap();
}

template <typename CalculatorType>
void default_write_status(CalculatorType const &mc_calculator, MethodLog &method_log) {
  std::ostream &sout = std::cout;
  methods::default_write_run_status(*mc_calculator.data, method_log, sout);
  sout << "  ParametricComposition="
       << mc_calculator.param_composition_calculator->per_unitcell()(0)
       << ", FormationEnergy=" << mc_calculator.formation_energy_calculator->per_unitcell()
       << std::endl;
  method_log.log.begin_lap();
}

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

  /// The device record of the sample being taken (set while sampling
  /// functions are evaluated during run).
  std::optional<mcg_sample> current_record;

  /// Observable of the current sample from the device's exact integers, with
  /// the reference's FP64 rounding (model.hh:266-295, 412-422;
  /// basic_semigrand_canonical.hh:165-174).
  double device_observable(int which) const {
    return detail::record_observable(*this, *current_record, which);
  }

  template <typename WriteStatusF = void (*)(SemiGrandCanonicalCalculator const &, MethodLog &)>
  std::shared_ptr<SemiGrandCanonicalData> run(
      state_type &_state, StateSamplingFunctionMap const &sampling_functions,
      jsonStateSamplingFunctionMap const &json_sampling_functions,
      CompletionCheckParams<BasicStatistics> const &completion_check_params,
      event_generator_type event_generator, int sample_period = 1,
      std::optional<MethodLog> method_log = std::nullopt,
      std::shared_ptr<engine_type> random_engine = nullptr,
      WriteStatusF write_status_f = default_write_status<SemiGrandCanonicalCalculator>) {
    // set-up (basic_semigrand_canonical.hh:427-462)
    state = &_state;
    conditions = std::make_shared<SemiGrandCanonicalConditions>(
        SemiGrandCanonicalConditions::from_values(state->conditions));
    const CountType N = state->configuration.n_variable_sites;
    potential.set_state(state, conditions);
    event_generator.set_state(state);
    data = std::make_shared<SemiGrandCanonicalData>(sampling_functions, json_sampling_functions,
                                                    N, completion_check_params);
    // the main loop (basic_occupation_metropolis.hh:354-425) on the device
    detail::run_sgc_on_device(*this, *data, sample_period, method_log, random_engine,
                              write_status_f);
    return data;
  }
};

template <typename CalculatorType>
StateSamplingFunction make_parametric_composition_f(std::shared_ptr<CalculatorType> mc) {
  if (mc == nullptr)
    throw std::runtime_error(
        "Error in parametric_composition sampling function: mc_calculator == nullptr");
  std::vector<Index> shape = {mc->system->param_composition_calculator.n_independent_compositions()};
  auto f = [mc]() -> VectorXd {
    if (mc->current_record) return VectorXd{mc->device_observable(MCG_OBS_PARAM_COMPOSITION)};
    return mc->param_composition_calculator->per_unitcell();
  };
  StateSamplingFunction s("param_composition", "Parametric composition", shape, f);
  s.device_observable = MCG_OBS_PARAM_COMPOSITION;
  return s;
}

template <typename CalculatorType>
StateSamplingFunction make_formation_energy_f(std::shared_ptr<CalculatorType> mc) {
  auto f = [mc]() -> VectorXd {
    if (mc == nullptr) throw std::runtime_error("Error in formation_energy sampling function");
    if (mc->current_record) return VectorXd{mc->device_observable(MCG_OBS_FORMATION_ENERGY)};
    return VectorXd{mc->formation_energy_calculator->per_unitcell()};
  };
  StateSamplingFunction s("formation_energy", "Intensive formation energy", {}, f);
  s.device_observable = MCG_OBS_FORMATION_ENERGY;
  return s;
}

template <typename CalculatorType>
StateSamplingFunction make_potential_energy_f(std::shared_ptr<CalculatorType> mc) {
  auto f = [mc]() -> VectorXd {
    if (mc == nullptr) throw std::runtime_error("Error in potential_energy sampling function");
    if (mc->current_record) return VectorXd{mc->device_observable(MCG_OBS_POTENTIAL_ENERGY)};
    return VectorXd{mc->potential.per_unitcell()};
  };
  StateSamplingFunction s("potential_energy", "Intensive potential energy", {}, f);
  s.device_observable = MCG_OBS_POTENTIAL_ENERGY;
  return s;
}

/// JSON text of the configuration: {"shape": [...], "occupation": [...]}.
template <typename CalculatorType>
jsonStateSamplingFunction make_configuration_json_f(std::shared_ptr<CalculatorType> mc) {
  auto f = [mc]() -> std::string {
    auto const &c = mc->state->configuration;
    std::ostringstream os;
    os << "{\"shape\": [" << c.shape[0] << ", " << c.shape[1] << "], \"occupation\": [";
    for (size_t i = 0; i < c.occupation().size(); ++i) os << (i ? ", " : "") << c.occupation()[i];
    os << "]}";
    return os.str();
  };
  return jsonStateSamplingFunction("configuration", "Configuration values", f);
}

}  // namespace basic_semigrand_canonical
}  // namespace ising_cpp
}  // namespace monte
}  // namespace CASM

#endif
