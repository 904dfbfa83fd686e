// Drop-in counterpart of libcasm-monte's kinetic Monte Carlo main loop
// (methods/kinetic_monte_carlo.hh:20-229): KMCData and kinetic_monte_carlo(),
// with the loop on the MI355X (KMC mode) for the square lattice gas of
// ising_cpp/kmc.hh and its event selector.
//
// Reference semantics kept (kinetic_monte_carlo.hh:152-226): occ_location.
// mol_size() steps per pass; every iteration checks completion, selects an
// event and its time increment, samples the fixtures due by count (state
// before the event, time of the last event), samples the fixtures due by time
// at each sample time the event's time reaches (kmc_data.time set to the
// sample time), then applies the event through the OccLocation with atom
// tracking, sets every fixture's time to the event's and counts a step and an
// accepted event.  Before each sample, kmc_data holds the fixture, its label,
// the selector's total rate at the selection and the atom positions; after
// it, prev_time and prev_atom_positions_cart of the fixture are updated.
// The device runs the iterations between the points where a fixture samples
// (by count, or an event's time reaching the next sample time) or a count
// cutoff or a periodic completion check falls, then stops with the next event
// selected and not applied (mcg_kmc_advance), so the samples see the
// reference's state, time and engine.
//
// Hop correlations (the CorrelationsData of the selected-event statistics) are
// recorded by the device when the selector's correlations_data_params is set,
// and left in its selected_event_data.
//
// Not supported: other event selectors and configurations, simulated-time
// cutoffs (min_time, max_time), selected-event histograms, unique atom ids.
#ifndef CASM_MI355X_monte_methods_kinetic_monte_carlo
#define CASM_MI355X_monte_methods_kinetic_monte_carlo

#include <algorithm>
#include <iostream>
#include <limits>
#include <map>
#include <string>

#include "casm/monte/gpu/Handle.hh"
#include "casm/monte/ising_cpp/kmc.hh"
#include "casm/monte/run_management/RunManager.hh"

namespace CASM {
namespace monte {

/// Data sampling functions can read (kinetic_monte_carlo.hh:22-95)
template <typename ConfigType, typename StatisticsType, typename EngineType>
struct KMCData {
  /// the current sampling fixture's label, set before it samples
  std::string sampling_fixture_label;
  /// the current sampling fixture, set before it samples
  SamplingFixture<ConfigType, StatisticsType, EngineType> const *sampling_fixture = nullptr;
  /// the total event rate at the sample
  double total_rate = 0.0;
  /// the time at the sample (by time: the sample time; by count: the time of
  /// the last event)
  double time = 0.0;
  /// time of the last sample, by fixture label
  std::map<std::string, double> prev_time;
  /// unique atom ids (not tracked here: empty)
  std::vector<Index> unique_atom_id;
  std::map<std::string, std::vector<Index>> prev_unique_atom_id;
  /// atom name index of each column of the position matrices (set by the caller)
  std::vector<Index> atom_name_index_list;
  /// atom positions (3 x atoms, as if there were no periodic boundary)
  MatrixXd atom_positions_cart;
  /// atom positions at the last sample, by fixture label
  std::map<std::string, MatrixXd> prev_atom_positions_cart;
};

/// kinetic_monte_carlo (kinetic_monte_carlo.hh:136-229) for the square lattice
/// gas: configurations of occupant indices (0 vacancy, 1 atom) of one
/// asymmetric unit whose occupants are {"Va", atom}, an OccLocation that
/// tracks atoms, and ising_cpp::LatticeGasKmcEventSelector.  get_event_f is
/// not called: the device applies the selected hops itself.
template <typename EventIDType, typename StatisticsType, typename EngineType,
          typename GetEventType>
void kinetic_monte_carlo(
    State<ising_cpp::IsingConfiguration> &state, OccLocation &occ_location,
    KMCData<ising_cpp::IsingConfiguration, StatisticsType, EngineType> &kmc_data,
    ising_cpp::LatticeGasKmcEventSelector &event_selector, GetEventType get_event_f,
    RunManager<ising_cpp::IsingConfiguration, StatisticsType, EngineType> &run_manager) {
  typedef ising_cpp::IsingConfiguration config_type;
  typedef SamplingFixture<config_type, StatisticsType, EngineType> fixture_type;
  auto &config = state.configuration;
  auto const &cv = occ_location.convert();
  const Index N = config.n_sites;
  if (event_selector.state() != &state)
    throw std::runtime_error("kinetic_monte_carlo (MI355X): selector not set to this state");
  if (cv.asym_size() != 1 || cv.occ_size(0) != 2 || cv.components_size(cv.species_index(0, 0)) != 0 ||
      cv.components_size(cv.species_index(0, 1)) != 1 || occ_location.mol_size() != N)
    throw std::runtime_error(
        "kinetic_monte_carlo (MI355X): one asymmetric unit with occupants {\"Va\", atom} on "
        "every site");
  if (occ_location.atom_size() != (Index)std::count(config.occupation().begin(),
                                                    config.occupation().end(), 1))
    throw std::runtime_error(
        "kinetic_monte_carlo (MI355X): the OccLocation must track atoms (update_atoms) and be "
        "initialized with the state");
  for (Index l = 0; l < N; ++l)
    if (config.occ(l) != 0 && config.occ(l) != 1)
      throw std::runtime_error("kinetic_monte_carlo (MI355X): occupations are occupant indices 0/1");
  bool stochastic = false;
  for (auto &f : run_manager.sampling_fixtures) {
    auto const &sp = f->params().sampling_params;
    auto const &cp = f->params().completion_check_params.cutoff_params;
    if (cp.min_time.has_value() || cp.max_time.has_value())
      throw std::runtime_error(
          "kinetic_monte_carlo (MI355X): simulated-time cutoffs (min_time, max_time) are not "
          "supported");
    if (sp.stochastic_sample_period) stochastic = true;
  }

  // the initial atoms: their translations are the device's displacements' origin
  std::vector<Atom> atoms0;
  for (Index a = 0; a < occ_location.atom_size(); ++a) atoms0.push_back(occ_location.atom(a));
  const Index n_atoms = (Index)atoms0.size();

  // kinetic_monte_carlo.hh:152-160
  kmc_data.time = 0.0;
  kmc_data.atom_positions_cart = occ_location.atom_positions_cart();
  kmc_data.prev_atom_positions_cart.clear();
  for (auto &f : run_manager.sampling_fixtures) {
    kmc_data.prev_time.emplace(f->label(), kmc_data.time);
    kmc_data.prev_atom_positions_cart.emplace(f->label(), kmc_data.atom_positions_cart);
  }

  mcg_chain_desc ch{};
  ch.shape[0] = config.shape[0];
  ch.shape[1] = config.shape[1];
  ch.shape[2] = 1;
  ch.temperature = event_selector.temperature;
  mcg_desc d{};
  d.mode = MCG_MODE_KMC;
  d.lattice = MCG_LATTICE_SQUARE;
  d.J1 = event_selector.J;
  d.kb = KB;
  d.sample_period = (int64_t)1 << 62;  // no records: samples read the state
  d.n_chains = 1;
  d.chains = &ch;
  gpu::Handle dev(d);
  dev.set_occupation(0, std::vector<int8_t>(config.occupation().begin(), config.occupation().end()));
  dev.load_engine(0, *event_selector.engine);
  // hop correlations, recorded by the device at every jumps_per_position_sample-th
  // jump of each atom
  auto const &hop_params = event_selector.correlations_data_params;
  if (hop_params.has_value()) {
    if (hop_params->stop_run_when_complete)
      throw std::runtime_error(
          "kinetic_monte_carlo (MI355X): stop_run_when_complete is not supported");
    if (!event_selector.selected_event_data)
      throw std::runtime_error(
          "kinetic_monte_carlo (MI355X): correlations_data_params without selected_event_data");
    dev.kmc_collect_hops(hop_params->jumps_per_position_sample,
                         hop_params->max_n_position_samples);
  }

  auto positions = [&]() {
    auto a = dev.atoms(0, n_atoms);
    MatrixXd R = MatrixXd::Zero(3, n_atoms);
    for (Index i = 0; i < n_atoms; ++i) {
      R(0, i) = (double)(atoms0[i].translation[0] + a[4 * i + 1]);
      R(1, i) = (double)(atoms0[i].translation[1] + a[4 * i + 2]);
      R(2, i) = (double)atoms0[i].translation[2];
    }
    return R;
  };
  auto sync_configuration = [&]() {
    auto occ = dev.occupation(0, N);
    config.set_occupation(std::vector<int>(occ.begin(), occ.end()));
  };
  double total_rate = 0.0;
  // kinetic_monte_carlo.hh:165-188
  auto pre_sample_action = [&](fixture_type &fixture, State<config_type> const &) {
    kmc_data.sampling_fixture_label = fixture.label();
    kmc_data.sampling_fixture = &fixture;
    kmc_data.atom_positions_cart = positions();
    kmc_data.total_rate = total_rate;
    if (fixture.params().sampling_params.sample_mode == SAMPLE_MODE::BY_TIME)
      kmc_data.time = fixture.next_sample_time();
  };
  auto post_sample_action = [&](fixture_type &fixture, State<config_type> const &) {
    kmc_data.prev_time[fixture.label()] = kmc_data.time;
    kmc_data.prev_atom_positions_cart[fixture.label()] = kmc_data.atom_positions_cart;
  };
  auto steps_of = [&](auto const &f, CountType c) {
    return f->params().sampling_params.sample_mode == SAMPLE_MODE::BY_STEP ? c : c * N;
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
  int64_t acc_prev = 0, g = 0;
  // the device's events since the last call, into the fixtures' counters
  auto account = [&]() {
    int64_t np = 0, step = 0, na = 0, nr = 0;
    double e = 0.0;
    dev.position(0, np, step, e);
    dev.counters(0, np, na, nr);
    const int64_t reached = np * N + step;
    for (auto &f : run_manager.sampling_fixtures)
      f->counter().increment_steps(reached - g, na - acc_prev, 0);
    acc_prev = na;
    g = reached;
    double t = 0.0, tp = 0.0;
    int64_t site = -1;
    dev.nfold_state(0, t, site, tp);
    run_manager.set_time(t);
    kmc_data.time = t;
    return std::make_pair(site, tp);
  };

  // kinetic_monte_carlo.hh:162-163
  run_manager.initialize(occ_location.mol_size());
  run_manager.update_next_sampling_fixture();
  const double inf = std::numeric_limits<double>::infinity();
  while (true) {
    if (run_manager.is_complete()) break;
    run_manager.write_status_if_due();
    bool due = false;
    for (auto &f : run_manager.sampling_fixtures)
      if (f->params().sampling_params.sample_mode != SAMPLE_MODE::BY_TIME &&
          f->counter().count == f->next_sample_count())
        due = true;
    std::pair<int64_t, double> pending;
    if (due) {
      // select the next event and leave it pending (:192-195)
      dev.kmc_advance(0, inf, true);
      pending = account();
    } else {
      // the iterations up to the next point where something can change
      int64_t target = g + (int64_t)1000000 * N;
      if (pending_checks()) target = g + 1;
      for (auto &f : run_manager.sampling_fixtures) {
        if (f->params().sampling_params.sample_mode != SAMPLE_MODE::BY_TIME)
          target = std::min<int64_t>(target, std::max<int64_t>(steps_of(f, f->next_sample_count()), g + 1));
        auto const &cp = f->params().completion_check_params.cutoff_params;
        const CountType count = f->counter().count;
        if (cp.min_count.has_value() && *cp.min_count > count)
          target = std::min<int64_t>(target, steps_of(f, *cp.min_count));
        if (cp.max_count.has_value() && *cp.max_count > count)
          target = std::min<int64_t>(target, steps_of(f, *cp.max_count));
      }
      const double stop = run_manager.next_sampling_fixture ? run_manager.next_sample_time : inf;
      dev.kmc_advance(target - g, stop, false);
      pending = account();
      if (pending.first < 0) continue;  // the next point is a count: check completion there
    }
    // the selected event: samples by count, then by time (:197-203)
    sync_configuration();
    total_rate = event_selector.total_rate();
    if (stochastic) dev.store_engine(0, *event_selector.engine);
    run_manager.sample_data_by_count_if_due(state, pre_sample_action, post_sample_action);
    run_manager.sample_data_by_time_if_due(pending.second, state, pre_sample_action,
                                           post_sample_action);
    if (stochastic) dev.load_engine(0, *event_selector.engine);
    // apply it, set the time, count the step (:205-215)
    dev.kmc_advance(1, inf, false);
    account();
  }
  sync_configuration();

  // the OccLocation the reference's applies leave: lists in the device's
  // order, Mol species and components, atoms with their translations and jumps
  {
    auto loc = dev.occ_location(0, N);
    auto a = dev.atoms(0, n_atoms);
    auto const &cl = occ_location.candidate_list();
    const Index va = cv.species_index(0, 0), at = cv.species_index(0, 1);
    const Index n_vac = N - n_atoms;
    std::vector<std::vector<Index>> lists(cl.size());
    for (Index i = 0; i < N; ++i)
      lists[cl.index(0, i < n_vac ? va : at)].push_back(occ_location.l_to_mol_id(loc[i]));
    std::vector<std::vector<Index>> component(N);
    std::vector<Atom> atoms(n_atoms);
    for (Index i = 0; i < n_atoms; ++i) {
      atoms[i].translation = atoms0[i].translation;
      atoms[i].translation[0] += a[4 * i + 1];
      atoms[i].translation[1] += a[4 * i + 2];
      atoms[i].n_jumps = atoms0[i].n_jumps + a[4 * i + 3];
      component[occ_location.l_to_mol_id(a[4 * i])].push_back(i);
    }
    for (Index l = 0; l < N; ++l)
      occ_location.mol(occ_location.l_to_mol_id(l)).species_index = cv.species_index(0, config.occ(l));
    occ_location.set_locations(lists);
    occ_location.set_atoms(atoms, component);
    kmc_data.atom_positions_cart = occ_location.atom_positions_cart();
  }
  if (hop_params.has_value()) {
    // the device's records as CorrelationsData (positions as if there were no
    // periodic boundary: the initial translation plus the displacement); the
    // complete samples are those every atom has
    const Index m = hop_params->max_n_position_samples;
    std::vector<int64_t> counts, rec;
    std::vector<double> times;
    dev.kmc_hops(0, n_atoms, m, counts, rec, times);
    CorrelationsData cd;
    cd.initialize(n_atoms, hop_params->jumps_per_position_sample, m,
                  hop_params->output_incomplete_samples);
    cd.stop_run_when_complete = false;
    CountType complete = n_atoms > 0 ? m : 0;
    for (Index a = 0; a < n_atoms; ++a) {
      cd.n_position_samples[a] = counts[a];
      complete = std::min<CountType>(complete, counts[a]);
      for (Index k = 0; k < counts[a]; ++k) {
        int64_t const *r = rec.data() + 4 * (k * n_atoms + a);
        cd.atom_positions_cart[k](0, a) = (double)(atoms0[a].translation[0] + r[0]);
        cd.atom_positions_cart[k](1, a) = (double)(atoms0[a].translation[1] + r[1]);
        cd.atom_positions_cart[k](2, a) = (double)atoms0[a].translation[2];
        cd.step(k, a) = r[2];
        cd.pass(k, a) = r[3];
        cd.time(k, a) = times[k * n_atoms + a];
      }
    }
    cd.n_complete_samples = complete;
    event_selector.selected_event_data->correlations_data = std::move(cd);
  }
  dev.store_engine(0, *event_selector.engine);
  (void)get_event_f;
  run_manager.finalize(state);
}

}  // namespace monte
}  // namespace CASM

#endif
