// Drop-in counterpart of libcasm-monte's RunManager-driven occupation
// Metropolis loop (methods/occupation_metropolis.hh:87-154) for the SGC Ising
// model with single-site flip proposals, with the loop on the MI355X.
//
// Reference semantics kept: the potential energy per unit cell is carried in
// state.properties["potential_energy"] as an FP64 running sum updated by
// fl(dE / n_unitcells) per accepted step (:103-106, :141) -- the device keeps
// this sum in the same order; fixtures sample by pass count; the run stops
// when the RunManager says so (global cutoff or all fixtures complete); the
// engine shared with the RunManager is advanced exactly as the reference
// advances it and written back at the end.
//
// The device runs the steps between the points where a fixture samples or a
// completion check can change its answer (detail::run_device_steps); fixtures
// sample by pass or, with the single-site SGC potential, by step; sampling by
// time has no meaning for Metropolis runs (there is no simulated time).
//
// Two entry points: the SGC Ising potential with single-site flips (EXACT
// mode), and the reference's generic signature with an OccLocation and
// propose_canonical_event or propose_semigrand_canonical_event (CANONICAL
// mode for one binary unit's pair swaps, OCCUPANT mode for the rest).
#ifndef CASM_MI355X_monte_methods_occupation_metropolis
#define CASM_MI355X_monte_methods_occupation_metropolis

#include <iostream>
#include <type_traits>

#include "casm/monte/ising_cpp/basic_semigrand_canonical.hh"
#include "casm/monte/ising_cpp/canonical.hh"
#include "casm/monte/run_management/RunManager.hh"

namespace CASM {
namespace monte {

/// SGC potential of the Ising model for occupation_metropolis: E_f - mu*N*x
/// per supercell (basic_semigrand_canonical.hh:165-169), bound to one state.
class IsingSemiGrandCanonicalPotential {
 public:
  IsingSemiGrandCanonicalPotential(double _J, bool _use_nlist) : J(_J), use_nlist(_use_nlist) {}
  double J;
  bool use_nlist;
  void set(State<ising_cpp::IsingConfiguration> const *s) { m_state = s; }
  State<ising_cpp::IsingConfiguration> const *state() const { return m_state; }
  double per_supercell() const {
    ising_cpp::IsingState st(m_state->configuration, m_state->conditions);
    ising_cpp::IsingFormationEnergy fe(J, 1, use_nlist, &st);
    ising_cpp::IsingParamComposition pc(&st);
    VectorXd mu = m_state->conditions.vector_values.at("exchange_potential");
    return fe.per_supercell() - mu.dot(pc.per_supercell());
  }

 private:
  State<ising_cpp::IsingConfiguration> const *m_state = nullptr;
};

namespace detail {

/// Drive a RunManager over device steps of chain 0 of `dev` (one pass = N
/// steps; g counts the steps since the start of the run).  The reference calls
/// the completion check before every step and samples after every step
/// (occupation_metropolis.hh:118-150).  Between two points where a fixture
/// samples or a count cutoff falls, the checks see the same samples and only
/// a changing count, so the device runs to the next such point in one launch:
///  - BY_PASS fixtures only: counts are constant within a pass, so the calls of
///    a pass after the first only repeat pending periodic checks, which the
///    host replays; the device runs whole passes;
///  - with BY_STEP fixtures (`steps_ok`, EXACT handles): a launch ends at any
///    step (mcg_run_steps); a periodic check still pending after a call is
///    due again at the next step's call, so the device then runs one step.
/// Counters, the running potential energy and, when a sampling function reads
/// it, the configuration (sync_configuration) come to the host at those points.
/// A fixture with a stochastic sampling period draws its next sample count
/// from the run's engine when it samples (SamplingFixture.hh:529-534, between
/// two steps of the walk): around such samples the walk's engine comes to the
/// host and goes back (engine_sync(true), engine_sync(false)).
template <typename ConfigType, typename StatisticsType, typename EngineType, typename SyncF,
          typename EngineSyncF>
void run_device_steps(State<ConfigType> &state, gpu::Handle &dev, CountType N, bool steps_ok,
                      RunManager<ConfigType, StatisticsType, EngineType> &run_manager,
                      SyncF sync_configuration, EngineSyncF engine_sync) {
  double &potential_energy_per_unitcell = state.properties.scalar_values["potential_energy"];
  run_manager.initialize(N);
  bool host_functions = false, by_step = false, stochastic = false;
  for (auto &f : run_manager.sampling_fixtures) {
    auto const &sp = f->params().sampling_params;
    if (sp.stochastic_sample_period) stochastic = true;
    if (sp.sample_mode == SAMPLE_MODE::BY_STEP) {
      if (!steps_ok)
        throw std::runtime_error(
            "occupation_metropolis (MI355X): SAMPLE_MODE::BY_STEP sampling runs with the "
            "single-site SGC potential only");
      by_step = true;
    } else if (sp.sample_mode != SAMPLE_MODE::BY_PASS) {
      throw std::runtime_error(
          "occupation_metropolis (MI355X): Metropolis runs sample by pass or by step "
          "(no simulated time)");
    }
    if (sp.do_sample_trajectory || !sp.json_sampler_names.empty()) host_functions = true;
    for (auto const &name : sp.sampler_names)
      if (f->params().sampling_functions.at(name).device_observable < 0) host_functions = true;
  }
  auto sample_if_due = [&]() {
    bool draws = false;
    if (stochastic)
      for (auto &f : run_manager.sampling_fixtures)
        if (f->params().sampling_params.stochastic_sample_period &&
            f->counter().count == f->next_sample_count())
          draws = true;
    if (draws) engine_sync(true);
    run_manager.sample_data_by_count_if_due(state);
    if (draws) engine_sync(false);
  };
  sample_if_due();
  // a fixture's count c in steps since the start
  auto steps_of = [&](auto const &f, CountType c) {
    return f->params().sampling_params.sample_mode == SAMPLE_MODE::BY_STEP ? c : c * N;
  };
  auto pending = [&]() {
    for (auto &f : run_manager.sampling_fixtures) {
      auto const &cc = f->completion_check();
      if (cc.results().is_complete) continue;
      if (cc.results().has_all_minimums_met &&
          get_n_samples(f->results().samplers) >= cc.next_check_at())
        return true;
    }
    return false;
  };

  CountType g = 0;
  int64_t acc_prev = 0, rej_prev = 0;
  while (true) {
    if (run_manager.is_complete()) break;
    if (!by_step)
      for (CountType c = 1; c < N && pending(); ++c) run_manager.is_complete();
    run_manager.write_status_if_due();

    // the earliest next point: the next step, or the next pass boundary when
    // every fixture counts passes
    const CountType next = by_step ? g + 1 : g + N;
    CountType target = g + 1000000 * N;
    if (pending()) target = next;
    for (auto &f : run_manager.sampling_fixtures) {
      const CountType count = f->counter().count;
      target = std::min(target, std::max(steps_of(f, f->next_sample_count()), next));
      auto const &cp = f->params().completion_check_params.cutoff_params;
      if (cp.min_count.has_value() && *cp.min_count > count)
        target = std::min(target, steps_of(f, *cp.min_count));
      if (cp.max_count.has_value() && *cp.max_count > count)
        target = std::min(target, steps_of(f, *cp.max_count));
    }
    const int rc = by_step ? dev.try_run_steps(target - g) : dev.try_run((target - g) / N);
    bool stopped = false;
    if (rc != MCG_OK) {
      const std::string msg = mcg_last_error();
      if (msg.rfind("Error in choose_", 0) != 0) gpu::check(rc);
      // occupation_metropolis.hh:118-127: the proposal found no events; the
      // loop logs the exception and stops after the steps made so far
      std::cerr << "\nError proposing event: " << msg << std::endl;
      stopped = true;
    }
    int64_t np = 0, step = 0, na = 0, nr = 0;
    double e_run = 0.0;
    dev.counters(0, np, na, nr);
    dev.position(0, np, step, e_run);
    const CountType reached = (CountType)np * N + (CountType)step;
    for (auto &f : run_manager.sampling_fixtures)
      f->counter().increment_steps(reached - g, na - acc_prev, nr - rej_prev);
    acc_prev = na;
    rej_prev = nr;
    g = reached;
    potential_energy_per_unitcell = e_run;
    if (stopped) break;
    if (host_functions) sync_configuration();
    sample_if_due();
  }
  sync_configuration();
}

}  // namespace detail

template <typename StatisticsType, typename EngineType>
void occupation_metropolis(
    State<ising_cpp::IsingConfiguration> &state, IsingSemiGrandCanonicalPotential &potential,
    RandomNumberGenerator<EngineType> &random_number_generator,
    RunManager<ising_cpp::IsingConfiguration, StatisticsType, EngineType> &run_manager) {
  if (potential.state() != &state)
    throw std::runtime_error(
        "Error in monte::occupation_metropolis: potential not set to correct state");
  auto &config = state.configuration;
  const double n_unitcells = static_cast<double>(config.n_sites);
  state.properties.scalar_values["potential_energy"] = 0.;
  state.properties.scalar_values["potential_energy"] = potential.per_supercell() / n_unitcells;

  mcg_chain_desc ch{};
  ch.shape[0] = config.shape[0];
  ch.shape[1] = config.shape[1];
  ch.shape[2] = 1;
  ch.temperature = state.conditions.scalar_values.at("temperature");
  ch.exchange_potential = state.conditions.vector_values.at("exchange_potential")(0);
  mcg_desc d{};
  d.mode = MCG_MODE_EXACT;
  d.lattice = MCG_LATTICE_SQUARE;
  d.use_nlist = potential.use_nlist ? 1 : 0;
  d.J1 = potential.J;
  d.kb = KB;
  d.sample_period = (int64_t)1 << 62;  // no records: samples read the chain's position
  d.n_chains = 1;
  d.chains = &ch;
  gpu::Handle dev(d);
  dev.set_occupation(0, std::vector<int8_t>(config.occupation().begin(), config.occupation().end()));
  dev.load_engine(0, *random_number_generator.engine);

  const CountType N = config.n_sites;
  detail::run_device_steps(
      state, dev, N, true, run_manager,
      [&]() {
        auto occ = dev.occupation(0, N);
        config.set_occupation(std::vector<int>(occ.begin(), occ.end()));
      },
      [&](bool to_host) {
        if (to_host)
          dev.store_engine(0, *random_number_generator.engine);
        else
          dev.load_engine(0, *random_number_generator.engine);
      });
  dev.store_engine(0, *random_number_generator.engine);
  run_manager.finalize(state);
}

namespace detail {

/// The OCCUPANT system of an OccLocation run: units, candidates in the
/// device's order (units with two occupants, occupant 0 then 1), the swap
/// types as candidate pairs and the exchange potentials.  Throws for what the
/// device does not run (units with more than two occupants, more than 8
/// units or species, more than 16 swap types).
struct OccupantSystem {
  std::vector<int64_t> occ_size, species, swaps;
  std::vector<double> mu;
  std::vector<int8_t> site_asym;
  std::vector<std::pair<Index, Index>> cand;  // device candidate -> (asym, species)
  mcg_occupant_desc desc{};

  OccupantSystem(Conversions const &cv, OccCandidateList const &cl,
                 std::vector<OccSwap> const &possible_swaps,
                 ising_cpp::IsingOccupantPotential const &potential, bool canonical,
                 Index n_sites, Index sites_per_unitcell) {
    const Index na = cv.asym_size(), ns = cv.species_size();
    if (na > 8 || ns > 8)
      throw std::runtime_error(
          "occupation_metropolis (MI355X): at most 8 asymmetric units and 8 species");
    for (Index b = 0; b < na; ++b) {
      if (cv.occ_size(b) > 2)
        throw std::runtime_error(
            "occupation_metropolis (MI355X): units allow at most two occupants");
      occ_size.push_back(cv.occ_size(b));
      for (Index o = 0; o < 2; ++o) species.push_back(o < cv.occ_size(b) ? cv.species_index(b, o) : 0);
      if (cv.occ_size(b) == 2)
        for (Index o = 0; o < 2; ++o) cand.push_back({b, cv.species_index(b, o)});
    }
    if ((Index)cl.size() != (Index)cand.size())
      throw std::runtime_error(
          "occupation_metropolis (MI355X): the candidate list must hold every (unit, species) "
          "of the units with two occupants");
    auto dev_cand = [&](OccCandidate const &c) -> int64_t {
      for (size_t i = 0; i < cand.size(); ++i)
        if (cand[i].first == c.asym && cand[i].second == c.species_index) return (int64_t)i;
      throw std::runtime_error("occupation_metropolis (MI355X): swap names an unknown candidate");
    };
    if (possible_swaps.empty() || possible_swaps.size() > 16)
      throw std::runtime_error("occupation_metropolis (MI355X): 1 to 16 swap types");
    for (auto const &sw : possible_swaps) {
      swaps.push_back(dev_cand(sw.cand_a));
      swaps.push_back(dev_cand(sw.cand_b));
    }
    if (potential.convert() && potential.convert() != &cv)
      throw std::runtime_error(
          "occupation_metropolis (MI355X): the potential's Conversions must be the OccLocation's");
    mu = potential.convert() ? potential.exchange_potential : std::vector<double>(ns, 0.0);
    for (Index l = 0; l < n_sites; ++l) site_asym.push_back((int8_t)cv.l_to_asym(l));
    desc.n_asym = na;
    desc.occ_size = occ_size.data();
    desc.species = species.data();
    desc.n_species = ns;
    desc.canonical = canonical ? 1 : 0;
    desc.n_swaps = (int64_t)possible_swaps.size();
    desc.swaps = swaps.data();
    desc.exchange_potential = mu.data();
    desc.sites_per_unitcell = sites_per_unitcell;
  }
};

}  // namespace detail

/// Runs through the reference's generic signature
/// (occupation_metropolis.hh:28-35, loop :87-153) over an OccLocation, for the
/// square-lattice Ising potential of occupant indices (ising_cpp/canonical.hh),
/// with propose_canonical_event or propose_semigrand_canonical_event and any
/// units of one or two occupants.  The chain runs on the MI355X: in CANONICAL
/// mode for one binary unit with its single canonical swap and no exchange
/// potentials, otherwise in OCCUPANT mode.  At the end the configuration, the
/// OccLocation lists (whose order decides later draws) and the engine are
/// where the reference's loop leaves them.
template <typename ProposeOccEventFuntionType, typename GeneratorType, typename StatisticsType,
          typename EngineType>
void occupation_metropolis(
    State<ising_cpp::IsingConfiguration> &state, OccLocation &occ_location,
    ising_cpp::IsingOccupantPotential &potential, std::vector<OccSwap> const &possible_swaps,
    ProposeOccEventFuntionType propose_event_f, GeneratorType &random_number_generator,
    RunManager<ising_cpp::IsingConfiguration, StatisticsType, EngineType> &run_manager) {
  if (potential.state() != &state)
    throw std::runtime_error(
        "Error in monte::occupation_metropolis: potential not set to correct state");
  typedef OccEvent &(*propose_f)(OccEvent &, OccLocation const &, std::vector<OccSwap> const &,
                                 GeneratorType &);
  bool canonical = false, semigrand = false;
  if constexpr (std::is_convertible_v<ProposeOccEventFuntionType, propose_f>) {
    const propose_f f = static_cast<propose_f>(propose_event_f);
    canonical = f == &propose_canonical_event<GeneratorType>;
    semigrand = f == &propose_semigrand_canonical_event<GeneratorType>;
  }
  if (!canonical && !semigrand)
    throw std::runtime_error(
        "occupation_metropolis (MI355X): the device runs propose_canonical_event and "
        "propose_semigrand_canonical_event");
  auto const &cv = occ_location.convert();
  auto const &cl = occ_location.candidate_list();
  auto &config = state.configuration;
  const CountType N = config.n_sites;
  if (cv.l_size() != N)
    throw std::runtime_error("occupation_metropolis (MI355X): Conversions and configuration sizes differ");
  for (Index l = 0; l < N; ++l)
    if (config.occ(l) < 0 || config.occ(l) >= cv.occ_size(cv.l_to_asym(l)))
      throw std::runtime_error(
          "occupation_metropolis (MI355X): configurations hold occupant indices of their units");

  // occupation_metropolis.hh:96-106
  const double n_unitcells = static_cast<double>(config.n_unitcells);
  state.properties.scalar_values["potential_energy"] = 0.;
  state.properties.scalar_values["potential_energy"] = potential.per_supercell() / n_unitcells;

  mcg_chain_desc ch{};
  ch.shape[0] = config.shape[0];
  ch.shape[1] = config.shape[1];
  ch.shape[2] = 1;
  ch.temperature = state.conditions.scalar_values.at("temperature");
  mcg_desc d{};
  d.lattice = MCG_LATTICE_SQUARE;
  d.use_nlist = potential.use_nlist ? 1 : 0;
  d.J1 = potential.J;
  d.kb = KB;
  d.sample_period = (int64_t)1 << 62;  // no records: samples read the chain's position
  d.n_chains = 1;
  d.chains = &ch;
  auto sync_engine = [&](gpu::Handle &dev) {
    return [&dev, &random_number_generator](bool to_host) {
      if (to_host)
        dev.store_engine(0, *random_number_generator.engine);
      else
        dev.load_engine(0, *random_number_generator.engine);
    };
  };

  const bool pair_mode = canonical && cv.asym_size() == 1 && cv.occ_size(0) == 2 &&
                         cl.size() == 2 && possible_swaps.size() == 1 &&
                         possible_swaps[0] == OccSwap(cl[0], cl[1]) && !potential.convert();
  if (pair_mode) {
    if (occ_location.cand_size(0) == 0 || occ_location.cand_size(1) == 0) {
      // OccEventProposal.hh:136-139 throws; occupation_metropolis.hh:121-127 logs and stops
      std::cerr << "\nError proposing event: Error in choose_canonical_swap: No events possible."
                << std::endl;
      run_manager.initialize(N);
      run_manager.sample_data_by_count_if_due(state);
      run_manager.finalize(state);
      return;
    }
    d.mode = MCG_MODE_CANONICAL;
    gpu::Handle dev(d);
    std::vector<int8_t> spins(N);
    for (Index l = 0; l < N; ++l) spins[l] = config.occ(l) ? 1 : -1;
    dev.set_occupation(0, spins);
    std::vector<int64_t> loc;
    for (auto const &v : occ_location.locations())
      for (Index id : v) loc.push_back(occ_location.mol(id).l);
    dev.set_occ_location(0, loc);
    dev.load_engine(0, *random_number_generator.engine);
    detail::run_device_steps(
        state, dev, N, false, run_manager,
        [&]() {
          auto occ = dev.occupation(0, N);
          for (Index l = 0; l < N; ++l) config.set_occ(l, occ[l] > 0 ? 1 : 0);
        },
        sync_engine(dev));
    auto dl = dev.occ_location(0, N);
    std::vector<std::vector<Index>> lists(2);
    const Index nA = occ_location.cand_size(0);
    for (Index i = 0; i < N; ++i) lists[i < nA ? 0 : 1].push_back(occ_location.l_to_mol_id(dl[i]));
    for (Index l = 0; l < N; ++l)
      occ_location.mol(occ_location.l_to_mol_id(l)).species_index = config.occ(l);
    occ_location.set_locations(lists);
    dev.store_engine(0, *random_number_generator.engine);
    run_manager.finalize(state);
    return;
  }

  if (N % config.n_unitcells != 0)
    throw std::runtime_error("occupation_metropolis (MI355X): sites per unit cell must be whole");
  detail::OccupantSystem sys(cv, cl, possible_swaps, potential, canonical, N,
                             N / config.n_unitcells);
  d.mode = MCG_MODE_OCCUPANT;
  d.occupant = &sys.desc;
  gpu::Handle dev(d);
  dev.set_site_asym(0, sys.site_asym);
  dev.set_occupation(0, std::vector<int8_t>(config.occupation().begin(), config.occupation().end()));
  // the lists in the device's candidate order, as sites
  const CountType n_mut = occ_location.mol_size();
  std::vector<int64_t> loc;
  for (auto const &c : sys.cand)
    for (Index i = 0; i < occ_location.cand_size(cl.index(c.first, c.second)); ++i)
      loc.push_back(occ_location.mol(occ_location.mol_id(cl.index(c.first, c.second), i)).l);
  dev.set_occ_location(0, loc);
  dev.load_engine(0, *random_number_generator.engine);

  detail::run_device_steps(
      state, dev, n_mut, true, run_manager,
      [&]() {
        auto occ = dev.occupation(0, N);
        config.set_occupation(std::vector<int>(occ.begin(), occ.end()));
      },
      sync_engine(dev));
  // Mol species from the final occupation, then the device's lists
  for (Index l = 0; l < N; ++l) {
    const Index id = occ_location.l_to_mol_id(l);
    if (id < n_mut) occ_location.mol(id).species_index = cv.species_index(cv.l_to_asym(l), config.occ(l));
  }
  auto dl = dev.occ_location(0, n_mut);
  std::vector<std::vector<Index>> lists(cl.size());
  {
    std::vector<Index> size(sys.cand.size(), 0);
    for (Index l = 0; l < N; ++l) {
      const Index id = occ_location.l_to_mol_id(l);
      if (id >= n_mut) continue;
      const Index b = cv.l_to_asym(l), sp = cv.species_index(b, config.occ(l));
      for (size_t i = 0; i < sys.cand.size(); ++i)
        if (sys.cand[i].first == b && sys.cand[i].second == sp) ++size[i];
    }
    Index k = 0;
    for (size_t i = 0; i < sys.cand.size(); ++i) {
      auto &v = lists[cl.index(sys.cand[i].first, sys.cand[i].second)];
      for (Index j = 0; j < size[i]; ++j) v.push_back(occ_location.l_to_mol_id(dl[k++]));
    }
  }
  occ_location.set_locations(lists);
  dev.store_engine(0, *random_number_generator.engine);
  run_manager.finalize(state);
}

}  // namespace monte
}  // namespace CASM

#endif
