// The device-driven main loop shared by the two SGC Ising calculators
// (ising_cpp/basic_semigrand_canonical.hh, whose loop is
// methods/basic_occupation_metropolis.hh:354-425, and
// ising_cpp/complete_semigrand_canonical.hh:575-680, which writes the same loop
// inline).  Both loops consume the same engine stream, take the same samples and
// call the completion check with the same arguments; the device runs the
// passes between the points where the check can change its answer:
//   - inside a pass the check sees a constant count and sample count, so only
//     its first call in a pass can return true; the other calls can only
//     repeat a periodic check on the same samples, which the host replays;
//   - between two samples nothing the check reads changes except the count;
//   - so the loop stops the device at the next sample at which a periodic
//     check, min/max_sample or a clock-time cutoff is due, or at the next
//     min/max_count.
// Sampling functions created by make_*_f are evaluated from the device's exact
// integer records with the reference's FP64 rounding; any other function is
// called after the host configuration has been refreshed (one sample per
// launch in that case).
#ifndef CASM_MI355X_monte_ising_cpp_sgc_device_run
#define CASM_MI355X_monte_ising_cpp_sgc_device_run

#include <filesystem>
#include <optional>

#include "casm/monte/MethodLog.hh"
#include "casm/monte/RandomNumberGenerator.hh"
#include "casm/monte/checks/CompletionCheck.hh"
#include "casm/monte/gpu/Handle.hh"
#include "casm/monte/ising_cpp/model.hh"

namespace CASM {
namespace monte {
namespace ising_cpp {
namespace detail {

/// Observable `which` (MCG_OBS_*) of a device record with the reference's FP64
/// rounding: E_f = fl(S1 * (-J)) or the device's row/column fold
/// (model.hh:259-295), N x = (N + M) / 2 (model.hh:412-422), E_sgc = E_f -
/// mu N x (basic_semigrand_canonical.hh:165-174), each per unit cell.
template <typename CalculatorType>
double record_observable(CalculatorType const &calc, mcg_sample const &r, int which) {
  const double N = static_cast<double>(calc.state->configuration.n_unitcells);
  const double Ef = calc.formation_energy_calculator->use_nlist()
                        ? static_cast<double>(r.S1) * -calc.formation_energy_calculator->J
                        : r.e_fold;
  if (which == MCG_OBS_FORMATION_ENERGY) return Ef / N;
  const double Nx = static_cast<double>(calc.state->configuration.n_sites + r.M) / 2.0;
  if (which == MCG_OBS_PARAM_COMPOSITION) return Nx / N;
  const double mu = calc.conditions->exchange_potential(0);
  return (Ef - mu * Nx) / N;
}

/// Run the SGC Metropolis loop of `calc` on the MI355X until its completion
/// check stops it.  `calc` has been set up (state, conditions, potential) and
/// `data` constructed; the calculator type provides `state`, `conditions`,
/// `formation_energy_calculator` (J, use_nlist) and `current_record`, the data
/// type the members of BasicOccupationMetropolisData.  On return the state's
/// configuration, the engine and data's counters hold the final values and
/// `write_status_f(calc, method_log)` has been called once more, as after the
/// reference's loop.
template <typename CalculatorType, typename DataType, typename EngineType,
          typename WriteStatusF>
void run_sgc_on_device(CalculatorType &calc, DataType &data, int sample_period,
                       std::optional<MethodLog> &method_log,
                       std::shared_ptr<EngineType> random_engine, WriteStatusF &&write_status_f) {
  if (sample_period < 1) throw std::runtime_error("sample_period must be >= 1");
  if (calc.conditions->exchange_potential.size() != 1)
    throw std::runtime_error("Ising SGC: exchange_potential must have one component");
  auto &state = *calc.state;
  const CountType N = data.n_steps_per_pass;

  // basic_occupation_metropolis.hh:364-373 (complete_semigrand_canonical.hh:622-631)
  RandomNumberGenerator<EngineType> rng(random_engine);
  if (!method_log.has_value()) {
    method_log = MethodLog();
    method_log->logfile_path = std::filesystem::current_path() / "status.json";
    method_log->log_frequency = 600.0;
  }
  method_log->log.restart_clock();
  method_log->log.begin_lap();

  // device chain
  mcg_chain_desc ch{};
  ch.shape[0] = state.configuration.shape[0];
  ch.shape[1] = state.configuration.shape[1];
  ch.shape[2] = 1;
  ch.temperature = calc.conditions->temperature;
  ch.exchange_potential = calc.conditions->exchange_potential(0);
  ch.seed = 0;
  mcg_desc d{};
  d.mode = MCG_MODE_EXACT;
  d.lattice = MCG_LATTICE_SQUARE;
  d.use_nlist = calc.formation_energy_calculator->use_nlist() ? 1 : 0;
  d.J1 = calc.formation_energy_calculator->J;
  d.J2 = 0.0;
  d.kb = KB;
  d.sample_period = sample_period;
  d.n_chains = 1;
  d.chains = &ch;
  d.options = 0;
  gpu::Handle dev(d);
  {
    std::vector<int8_t> occ(state.configuration.occupation().begin(),
                            state.configuration.occupation().end());
    dev.set_occupation(0, occ);
  }
  dev.load_engine(0, *rng.engine);
  auto sync_configuration = [&]() {
    auto occ = dev.occupation(0, state.configuration.n_sites);
    state.configuration.set_occupation(std::vector<int>(occ.begin(), occ.end()));
  };

  bool host_functions = !data.json_sampling_functions.empty();
  for (auto const &p : data.sampling_functions)
    if (p.second.device_observable < 0) host_functions = true;
  auto const &cc_params = data.completion_check.params();
  const bool clock_cutoffs = cc_params.cutoff_params.min_clocktime.has_value() ||
                             cc_params.cutoff_params.max_clocktime.has_value();
  Log &log = method_log->log;
  auto &cc = data.completion_check;

  while (true) {
    // the first completion check of the current pass
    if (cc.is_complete(data.samplers, data.sample_weight, data.n_pass, log)) break;
    // the other calls of this pass: repeat periodic checks on the same samples
    for (CountType c = 1; c < N; ++c) {
      CountType ns = get_n_samples(data.samplers);
      if (!(cc.results().has_all_minimums_met && ns >= cc.next_check_at())) break;
      cc.is_complete(data.samplers, data.sample_weight, data.n_pass, log);
    }
    // next pass at which a completion-check call can answer differently
    const CountType ns = get_n_samples(data.samplers);
    const CountType p0 = data.n_pass;
    auto pass_of_sample = [&](CountType s) { return s * (CountType)sample_period; };
    CountType s_next = ns + 1;
    CountType target;
    if (cc.results().has_all_minimums_met && cc.next_check_at() <= ns) {
      // periodic checks still pending after a whole pass of calls: replay the
      // next pass's calls too
      target = p0 + 1;
    } else if (host_functions || clock_cutoffs) {
      target = pass_of_sample(s_next);
    } else {
      CountType s_t = std::max(s_next, cc.next_check_at());
      auto const &cp = cc_params.cutoff_params;
      if (cp.min_sample.has_value() && *cp.min_sample > ns) s_t = std::min(s_t, *cp.min_sample);
      if (cp.max_sample.has_value() && *cp.max_sample > ns) s_t = std::min(s_t, *cp.max_sample);
      target = pass_of_sample(s_t);
    }
    auto const &cp = cc_params.cutoff_params;
    if (cp.min_count.has_value() && *cp.min_count > p0) target = std::min(target, *cp.min_count);
    if (cp.max_count.has_value() && *cp.max_count > p0) target = std::min(target, *cp.max_count);
    target = std::max(target, p0 + 1);
    target = std::min(target, p0 + (CountType)1000000);

    dev.run(target - p0);
    const Index n_new = dev.n_samples(0) - ns;
    std::vector<mcg_sample> recs = dev.samples(0, ns, n_new);
    if (host_functions) sync_configuration();
    for (auto const &rec : recs) {
      data.n_pass = rec.pass;
      calc.current_record = rec;
      for (auto const &p : data.sampling_functions)
        data.samplers.at(p.second.name)->push_back(p.second());
      for (auto const &p : data.json_sampling_functions)
        data.json_samplers.at(p.second.name).values.push_back(p.second());
      if (method_log->log_frequency.has_value() &&
          method_log->log.lap_time() >= method_log->log_frequency.value()) {
        sync_configuration();
        write_status_f(calc, *method_log);
      }
    }
    calc.current_record.reset();
    data.n_pass = target;
  }

  // final state: configuration, engine and counters
  sync_configuration();
  dev.store_engine(0, *rng.engine);
  int64_t n_pass = 0, n_acc = 0, n_rej = 0;
  dev.counters(0, n_pass, n_acc, n_rej);
  data.n_accept = n_acc;
  data.n_reject = n_rej;
  write_status_f(calc, *method_log);
}

}  // namespace detail
}  // namespace ising_cpp
}  // namespace monte
}  // namespace CASM

#endif
