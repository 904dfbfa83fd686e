// Drop-in counterpart of libcasm-monte's N-fold way main loop
// (methods/nfold.hh:20-152) for the binary square-lattice Ising model in the
// semi-grand canonical ensemble, with the loop on the MI355X (NFOLD mode).
//
// The reference's loop is generic over an event selector (an external
// rejection-free selector in practice).  Here the selector is
// ising_cpp::IsingNfoldEventSelector: events are the single-site flips (event
// id = linear site index), the rate of a flip is its Metropolis probability
// min(1, exp(-dE / kT)) with dE the Ising energy change minus
// exchange_potential times the change of the spin sum, and the event is drawn
// by the class-partitioned selector of the device kernel from the selector's
// engine.  Configurations hold occupant indices 0/1 (occupant index 1 is spin
// +1), as OccLocation requires.
//
// Reference semantics kept (nfold.hh:83-150): one step per event and
// occ_location.mol_size() steps per pass; a fixture due at count c samples the
// state before the first event of pass c, after that event has been selected,
// with weight = its time increment and expected_acceptance_rate = total_rate /
// n_events_possible (:93-110); every event is counted as accepted; the
// completion check runs before every event, so a check that stops the run
// right after a sample stops it after the sampled event is applied, as the
// reference's loop does; the fixtures' time is the time of the last event.
// Between those points the device runs whole passes.
//
// Fixtures that sample by time (SAMPLE_MODE::BY_TIME) sample with weight 1 at
// every sample time an event's time reaches, before that event is applied
// (nfold.hh:93-110, :128-130); with such a fixture the device runs event by
// event between the points where anything can change and stops with the event
// that reaches the next sample time selected (mcg_kmc_advance).
//
// Not supported: SAMPLE_MODE::BY_STEP fixtures, stochastic sample periods,
// simulated-time cutoffs, and selected-event statistics.  At the end occ_location is rebuilt from the final occupation:
// its candidate lists hold the same sites as the reference's, in site order
// rather than in the order the reference's applies leave them (only a later
// draw from those lists would notice).
#ifndef CASM_MI355X_monte_methods_nfold
#define CASM_MI355X_monte_methods_nfold

#include <algorithm>
#include <limits>
#include <memory>
#include <string>

#include "casm/monte/events/OccLocation.hh"
#include "casm/monte/gpu/Handle.hh"
#include "casm/monte/ising_cpp/model.hh"
#include "casm/monte/run_management/RunManager.hh"

namespace CASM {
namespace monte {

/// Data that sampling functions can read (nfold.hh:20-42)
template <typename ConfigType, typename StatisticsType, typename EngineType>
struct NfoldData {
  /// set to the current sampling fixture's label before it samples
  std::string sampling_fixture_label;
  /// set to the current sampling fixture before it samples
  monte::SamplingFixture<ConfigType, StatisticsType, EngineType> const *sampling_fixture = nullptr;
  /// number of events that could be selected at any time; set by the caller
  double n_events_possible = 0.0;
  /// total_rate / n_events_possible at the sample
  double expected_acceptance_rate = 0.0;
};

namespace ising_cpp {

/// The N-fold way event selector of the SGC Ising model for nfold(): single-
/// site flips with Metropolis rates (see the top of this file), drawn from
/// `engine`, which nfold() advances and writes back.  Temperature and
/// exchange_potential come from the state's conditions.
template <typename EngineType>
class IsingNfoldEventSelector {
 public:
  IsingNfoldEventSelector(double _J, std::shared_ptr<EngineType> _engine, bool _use_nlist = true)
      : J(_J), use_nlist(_use_nlist), engine(_engine) {
    if (!engine) throw std::runtime_error("IsingNfoldEventSelector: engine==nullptr");
  }
  double J;
  bool use_nlist;
  std::shared_ptr<EngineType> engine;
};

}  // namespace ising_cpp

template <typename EventIDType, typename GetEventType, typename StatisticsType,
          typename EngineType>
void nfold(State<ising_cpp::IsingConfiguration> &state, OccLocation &occ_location,
           NfoldData<ising_cpp::IsingConfiguration, StatisticsType, EngineType> &nfold_data,
           ising_cpp::IsingNfoldEventSelector<EngineType> &event_selector,
           GetEventType get_event_f,
           RunManager<ising_cpp::IsingConfiguration, StatisticsType, EngineType> &run_manager) {
  typedef ising_cpp::IsingConfiguration config_type;
  typedef SamplingFixture<config_type, StatisticsType, EngineType> fixture_type;
  auto &config = state.configuration;
  const CountType N = config.n_sites;
  auto const &cv = occ_location.convert();
  if (cv.asym_size() != 1 || cv.occ_size(0) != 2 || occ_location.candidate_list().size() != 2)
    throw std::runtime_error("nfold (MI355X): needs one binary asymmetric unit");
  if (occ_location.mol_size() != N)
    throw std::runtime_error("nfold (MI355X): every site must mutate");
  for (Index l = 0; l < N; ++l)
    if (config.occ(l) != 0 && config.occ(l) != 1)
      throw std::runtime_error("nfold (MI355X): configurations hold occupant indices 0/1");
  bool by_time = false;
  for (auto &f : run_manager.sampling_fixtures) {
    auto const &sp = f->params().sampling_params;
    if (sp.sample_mode == SAMPLE_MODE::BY_STEP)
      throw std::runtime_error(
          "nfold (MI355X): SAMPLE_MODE::BY_PASS and BY_TIME sampling are supported, BY_STEP is not");
    if (sp.sample_mode == SAMPLE_MODE::BY_TIME) by_time = true;
    if (sp.stochastic_sample_period)
      throw std::runtime_error("nfold (MI355X): stochastic sample periods are not supported");
    // the reference checks simulated-time cutoffs before every event
    // (SamplingFixture.hh:272-275); the device stops only between passes
    auto const &cp = f->params().completion_check_params.cutoff_params;
    if (sp.do_sample_time && (cp.min_time.has_value() || cp.max_time.has_value()))
      throw std::runtime_error(
          "nfold (MI355X): simulated-time cutoffs (min_time, max_time) are not supported");
  }

  // nfold.hh:118-119
  run_manager.initialize(occ_location.mol_size());

  mcg_chain_desc ch{};
  ch.shape[0] = config.shape[0];
  ch.shape[1] = config.shape[1];
  ch.shape[2] = 1;
  ch.temperature = state.conditions.scalar_values.at("temperature");
  ch.exchange_potential = state.conditions.vector_values.at("exchange_potential")(0);
  mcg_desc d{};
  d.mode = MCG_MODE_NFOLD;
  d.lattice = MCG_LATTICE_SQUARE;
  d.use_nlist = event_selector.use_nlist ? 1 : 0;
  d.J1 = event_selector.J;
  d.kb = KB;
  d.sample_period = (int64_t)1 << 62;  // records only when asked for (mcg_run of 0 passes)
  d.n_chains = 1;
  d.chains = &ch;
  gpu::Handle dev(d);
  std::vector<int8_t> spins(N);
  for (Index l = 0; l < N; ++l) spins[l] = config.occ(l) ? 1 : -1;
  dev.set_occupation(0, spins);
  dev.load_engine(0, *event_selector.engine);

  auto sync_configuration = [&]() {
    auto occ = dev.occupation(0, N);
    for (Index l = 0; l < N; ++l) config.set_occ(l, occ[l] > 0 ? 1 : 0);
  };

  double total_rate = 0.0, time_increment = 0.0;
  auto pre_sample_action = [&](fixture_type &fixture, State<config_type> const &) {
    nfold_data.sampling_fixture_label = fixture.label();
    nfold_data.sampling_fixture = &fixture;
    nfold_data.expected_acceptance_rate = total_rate / nfold_data.n_events_possible;
    fixture.push_back_sample_weight(time_increment);
  };
  auto pending_checks = [&]() {
    for (auto &f : run_manager.sampling_fixtures) {
      auto const &cc = f->completion_check();
      if (cc.results().is_complete) continue;
      if (cc.results().has_all_minimums_met &&
          get_n_samples(f->results().samplers) >= cc.next_check_at())
        return true;
    }
    return false;
  };

  if (by_time) {
    // Event by event (nfold.hh:118-149): the device runs the iterations up to
    // the next point where something can change (a count due, a count cutoff,
    // a periodic check, or an event whose time reaches the next sample time)
    // and stops there with that event selected; the fixtures sample by count,
    // then by time, and the event is applied.
    auto steps_of = [&](CountType c) { return (int64_t)c * (int64_t)N; };
    int64_t acc = 0, g = 0;
    double time = 0.0;
    auto account = [&]() {
      int64_t np = 0, step = 0, na = 0, nr = 0;
      double rate = 0.0, dt = 0.0;
      dev.position(0, np, step, rate);
      dev.counters(0, np, na, nr);
      const int64_t reached = np * (int64_t)N + step;
      for (auto &f : run_manager.sampling_fixtures) f->counter().increment_steps(reached - g, na - acc, 0);
      acc = na;
      g = reached;
      int64_t site = -1;
      dev.nfold_state(0, time, site, dt);
      run_manager.set_time(time);
      total_rate = rate;
      time_increment = dt;
      return site;
    };
    auto pre_sample_time = [&](fixture_type &fixture, State<config_type> const &) {
      nfold_data.sampling_fixture_label = fixture.label();
      nfold_data.sampling_fixture = &fixture;
      nfold_data.expected_acceptance_rate = total_rate / nfold_data.n_events_possible;
      if (fixture.params().sampling_params.sample_mode == SAMPLE_MODE::BY_TIME)
        fixture.push_back_sample_weight(1.0);
      else
        fixture.push_back_sample_weight(time_increment);
    };
    typename RunManager<config_type, StatisticsType, EngineType>::NullAction post_sample;
    run_manager.update_next_sampling_fixture();
    const double inf = std::numeric_limits<double>::infinity();
    while (!run_manager.is_complete()) {
      run_manager.write_status_if_due();
      bool due = false;
      for (auto &f : run_manager.sampling_fixtures)
        if (f->params().sampling_params.sample_mode != SAMPLE_MODE::BY_TIME &&
            f->counter().count == f->next_sample_count())
          due = true;
      int64_t site = -1;
      if (due) {
        dev.kmc_advance(0, inf, true);  // select the event, leave it pending
        site = account();
      } else {
        int64_t target = g + (int64_t)1000000 * (int64_t)N;
        if (pending_checks()) target = g + 1;
        for (auto &f : run_manager.sampling_fixtures) {
          if (f->params().sampling_params.sample_mode != SAMPLE_MODE::BY_TIME)
            target = std::min<int64_t>(target, std::max<int64_t>(steps_of(f->next_sample_count()), g + 1));
          auto const &cp = f->params().completion_check_params.cutoff_params;
          const CountType count = f->counter().count;
          if (cp.min_count.has_value() && *cp.min_count > count)
            target = std::min<int64_t>(target, steps_of(*cp.min_count));
          if (cp.max_count.has_value() && *cp.max_count > count)
            target = std::min<int64_t>(target, steps_of(*cp.max_count));
        }
        const double stop = run_manager.next_sampling_fixture ? run_manager.next_sample_time : inf;
        dev.kmc_advance(target - g, stop, false);
        site = account();
        if (site < 0) continue;  // a count point: the completion check comes first
      }
      // the selected event: samples by count, then by time (nfold.hh:121-126)
      const double event_time = time + time_increment;
      sync_configuration();
      run_manager.sample_data_by_count_if_due(state, pre_sample_time, post_sample);
      run_manager.sample_data_by_time_if_due(event_time, state, pre_sample_time, post_sample);
      // apply it, set the time, count the step (:128-141)
      dev.kmc_advance(1, inf, false);
      account();
    }
    sync_configuration();
    {
      std::vector<int> occ(config.occupation().begin(), config.occupation().end());
      occ_location.initialize(occ);
    }
    dev.store_engine(0, *event_selector.engine);
    (void)get_event_f;
    run_manager.finalize(state);
    return;
  }

  CountType p = 0;
  int64_t acc_prev = 0;
  bool stopped_in_pass = false;
  while (!run_manager.is_complete()) {
    run_manager.write_status_if_due();
    bool due = false;
    for (auto &f : run_manager.sampling_fixtures) due |= f->counter().count == f->next_sample_count();
    if (due) {
      // nfold.hh:121-126: the event is selected, then the due fixtures sample;
      // the record holds the state before the pass's first event, that
      // event's increment and the total rate
      dev.run(0);
      const mcg_sample cur = dev.samples(0, dev.n_samples(0) - 1, 1)[0];
      dev.clear_samples();
      total_rate = cur.e_run;
      time_increment = cur.weight;
      sync_configuration();
      run_manager.sample_data_by_count_if_due(state, pre_sample_action);
      // the check after the first event of this pass (nfold.hh:118, :132-141);
      // with one site that event ends the pass, and the check at the next
      // pass boundary (the loop's own) is that check
      if (N > 1) {
        double t = 0.0, dt = 0.0;
        int64_t site = -1;
        dev.nfold_state(0, t, site, dt);
        if (site < 0) throw std::runtime_error("nfold (MI355X): no pending event at a sample");
        // that check sees the time after the event (nfold.hh:137-144: the
        // event time is set before the next loop test); time cutoffs are
        // refused above, so only the results' recorded time depends on it
        run_manager.set_time(t + dt);
        if (run_manager.is_complete()) {
          std::vector<int> occ(config.occupation().begin(), config.occupation().end());
          occ_location.initialize(occ);
          run_manager.increment_n_accept();
          occ_location.apply(get_event_f(static_cast<EventIDType>(site)), occ);
          config.set_occupation(occ);
          run_manager.increment_step();
          stopped_in_pass = true;
          break;
        }
        run_manager.set_time(t);
      }
    }
    // the remaining checks of this pass only repeat periodic checks on the same samples
    for (CountType c = 2; c < N && pending_checks(); ++c) run_manager.is_complete();
    if (run_manager.sampling_fixtures.empty()) break;

    CountType target = p + 100000;
    // one site: the check after this pass's (single) event is the check at the
    // next pass boundary, and the sample just taken may already complete the
    // run (max_sample, or min_sample met with a check due), so run one pass
    if (pending_checks() || (due && N == 1)) target = p + 1;
    for (auto &f : run_manager.sampling_fixtures) {
      target = std::min(target, std::max(f->next_sample_count(), p + 1));
      auto const &cp = f->params().completion_check_params.cutoff_params;
      if (cp.min_count.has_value() && *cp.min_count > p) target = std::min(target, *cp.min_count);
      if (cp.max_count.has_value() && *cp.max_count > p) target = std::min(target, *cp.max_count);
    }
    dev.run(target - p);
    int64_t np = 0, na = 0, nr = 0, site = -1;
    double t = 0.0, dt = 0.0;
    dev.counters(0, np, na, nr);
    dev.nfold_state(0, t, site, dt);
    for (auto &f : run_manager.sampling_fixtures) {
      f->counter().increment_passes(target - p, na - acc_prev, 0);
      f->set_time(t);
    }
    acc_prev = na;
    p = target;
  }
  if (!stopped_in_pass) sync_configuration();
  {
    std::vector<int> occ(config.occupation().begin(), config.occupation().end());
    occ_location.initialize(occ);
  }
  dev.store_engine(0, *event_selector.engine);
  run_manager.finalize(state);
}

}  // namespace monte
}  // namespace CASM

#endif
