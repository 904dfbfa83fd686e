// Tests of the C++ drop-in classes (include/casm/monte/...) against the oracle
// (oracle/refcpu.hh).
//
//   dropin_tests --cpu : host logic only (completion check, statistics,
//                        equilibration, sampler) on synthetic series;
//   dropin_tests --gpu : SemiGrandCanonicalCalculator::run, occupation_metropolis
//                        and nfold on the MI355X vs the oracle's loops (bit for
//                        bit; N-fold way times to the rounding of log1p).
// Prints one line per case and "ALL OK" at the end; exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <numeric>

#include <unistd.h>

#include "casm/monte/ising_cpp/basic_semigrand_canonical.hh"
#include "casm/monte/ising_cpp/complete_semigrand_canonical.hh"
#include "casm/monte/methods/kinetic_monte_carlo.hh"
#include "casm/monte/methods/nfold.hh"
#include "casm/monte/methods/occupation_metropolis.hh"
#include "casm/monte/run_management/io/json/jsonResultsIO.hh"
#include "casm/monte/sampling/io/json/SelectedEventFunctions_json_io.hh"
#include "refcpu.hh"

using namespace CASM;
using namespace CASM::monte;
namespace sgc = CASM::monte::ising_cpp::basic_semigrand_canonical;

static int g_fail = 0;
#define EXPECT(cond, msg)                                                  \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::cout << "FAIL: " << msg << " (" << #cond << ")" << std::endl; \
      ++g_fail;                                                            \
    }                                                                      \
  } while (0)

static bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// ---------------------------------------------------------------------------
// CPU: host logic vs oracle
// ---------------------------------------------------------------------------
static std::vector<double> synthetic_series(int kind, int n, uint64_t seed) {
  std::mt19937_64 e(seed);
  std::normal_distribution<double> g(0.0, 1.0);
  std::vector<double> x(n);
  double ar = 0.0;
  for (int i = 0; i < n; ++i) {
    double eps = g(e);
    if (kind == 0) x[i] = 1.0 + 0.1 * eps;                       // iid
    if (kind == 1) x[i] = (ar = 0.9 * ar + eps) + 5.0;           // AR(1) rho=0.9
    if (kind == 2) x[i] = 3.0 * std::exp(-i / 50.0) + 0.05 * eps;  // relaxation + noise
    if (kind == 3) x[i] = 2.5;                                   // constant
  }
  return x;
}

static void cpu_tests() {
  // statistics and equilibration: identical arithmetic to the oracle
  for (int kind = 0; kind < 4; ++kind)
    for (int n : {1, 2, 7, 100, 1001}) {
      auto x = synthetic_series(kind, n, 100 + kind);
      VectorXd v(x);
      auto a = BasicStatisticsCalculator()(v);
      auto b = refcpu::basic_statistics(x);
      EXPECT(same_bits(a.mean, b.mean) && same_bits(a.calculated_precision, b.calculated_precision),
             "basic statistics kind " << kind << " n " << n);
      for (double prec : {1e-3, 0.05, 1.0}) {
        auto ea = _default_equilibration_check(v, prec);
        auto eb = refcpu::default_equilibration_check_impl(x, prec);
        EXPECT(ea.is_equilibrated == eb.is_equilibrated &&
                   ea.N_samples_for_equilibration == eb.N_samples_for_equilibration,
               "equilibration kind " << kind << " n " << n << " prec " << prec);
      }
    }
  // closed forms: constant series -> factor 1 and precision 0
  {
    VectorXd c(std::vector<double>(50, 2.5));
    EXPECT(autocorrelation_factor(c) == 1.0, "constant series autocorrelation factor");
    EXPECT(BasicStatisticsCalculator()(c).calculated_precision == 0.0, "constant precision");
  }
  // test_CompletionCheck.py:5-37: max_count = 12 stops at exactly 12
  {
    CompletionCheckParams<BasicStatistics> p;
    p.cutoff_params.max_count = 12;
    CompletionCheck<BasicStatistics> cc(p);
    SamplerMap samplers;
    samplers["e"] = std::make_shared<Sampler>(std::vector<Index>{});
    samplers["x"] = std::make_shared<Sampler>(std::vector<Index>{3});
    Sampler weight({});
    Log log;
    CountType n_steps = 0;
    while (!cc.is_complete(samplers, weight, n_steps, log)) {
      n_steps += 1;
      if (n_steps % 10 == 0) {
        samplers["e"]->push_back(VectorXd{0.0});
        samplers["x"]->push_back(VectorXd{0.0, 0.0, 0.0});
      }
    }
    EXPECT(n_steps == 12, "max_count KAT");
  }
  // completion check on identical sample streams: same decisions step by step
  for (int kind = 0; kind < 3; ++kind) {
    CompletionCheckParams<BasicStatistics> p;
    p.cutoff_params.min_sample = 100;
    p.check_begin = 100;
    p.check_period = 10;
    p.requested_precision[SamplerComponent("e", 0, "0")] = RequestedPrecision::abs(0.01);
    CompletionCheck<BasicStatistics> cc(p);
    refcpu::CompletionCheckParams q;
    q.cutoff_params.min_sample = 100;
    q.check_begin = 100;
    q.check_period = 10;
    refcpu::RequestedPrecision rq;
    rq.abs_convergence_is_required = true;
    rq.abs_precision = 0.01;
    q.requested_precision[refcpu::SamplerComponent{"e", 0}] = rq;
    refcpu::CompletionCheck cr(q);
    SamplerMap sa;
    sa["e"] = std::make_shared<Sampler>(std::vector<Index>{});
    refcpu::SamplerMap sb;
    sb["e"] = std::make_shared<refcpu::Sampler>(1);
    Sampler weight({});
    Log log;
    auto x = synthetic_series(kind, 20000, 7 + kind);
    bool da = false, db = false;
    CountType i = 0;
    for (; i < (CountType)x.size(); ++i) {
      da = cc.is_complete(sa, weight, i, log);
      db = cr.is_complete(sb, i);
      if (da != db || da) break;
      sa["e"]->push_back(VectorXd{x[i]});
      sb["e"]->push_back({x[i]});
    }
    EXPECT(da == db && cc.n_checks() == cr.n_checks(),
           "completion decisions kind " << kind << " at " << i);
  }
  std::cout << "cpu: completion check, statistics, equilibration, sampler: done" << std::endl;
}

// ---------------------------------------------------------------------------
// GPU: SemiGrandCanonicalCalculator::run vs the oracle
// ---------------------------------------------------------------------------
typedef ising_cpp::IsingSystem system_type;
typedef sgc::SemiGrandCanonicalEventGenerator<std::mt19937_64> event_generator_type;
typedef sgc::SemiGrandCanonicalCalculator<system_type, event_generator_type> calculator_type;

struct Case {
  int rows, cols;
  bool use_nlist;
  double T, mu;
  uint64_t seed;
  int sample_period;
  long max_count;     // -1: none
  double prec;        // requested abs precision of param_composition and potential_energy (<=0: none)
  long min_sample;    // -1: none
  bool host_function; // add a host-evaluated sampling function
};

static void gpu_case(Case const &c) {
  auto system = std::make_shared<system_type>(ising_cpp::IsingFormationEnergy(0.1, 1, c.use_nlist),
                                              ising_cpp::IsingParamComposition());
  auto calc = std::make_shared<calculator_type>(system);
  StateSamplingFunctionMap functions;
  for (auto f : {sgc::make_parametric_composition_f(calc), sgc::make_formation_energy_f(calc),
                 sgc::make_potential_energy_f(calc)})
    functions.emplace(f.name, f);
  if (c.host_function) {
    // magnetisation from the host configuration (not a device observable)
    StateSamplingFunction m("magnetisation", "sum of spins", {}, [calc]() -> VectorXd {
      double s = 0;
      for (int v : calc->state->configuration.occupation()) s += v;
      return VectorXd{s};
    });
    functions.emplace(m.name, m);
  }
  CompletionCheckParams<BasicStatistics> ccp;
  if (c.max_count > 0) ccp.cutoff_params.max_count = c.max_count;
  if (c.min_sample > 0) ccp.cutoff_params.min_sample = c.min_sample;
  ccp.check_begin = 100;
  ccp.check_period = 10;
  if (c.prec > 0) converge(functions, ccp).set_precision("potential_energy", c.prec).set_precision("param_composition", c.prec);

  ValueMap cond;
  cond.scalar_values["temperature"] = c.T;
  cond.vector_values["exchange_potential"] = VectorXd{c.mu};
  ising_cpp::IsingState state(ising_cpp::IsingConfiguration({c.rows, c.cols}, 1), cond);
  std::mt19937_64 e0(c.seed);
  // a few random flips so the start is not uniform
  for (int k = 0; k < 5; ++k) state.configuration.set_occ((k * 7919) % (c.rows * c.cols), -1);
  auto engine = std::make_shared<std::mt19937_64>(c.seed);
  MethodLog mlog;
  auto data = calc->run(state, functions, jsonStateSamplingFunctionMap(), ccp,
                        event_generator_type(), c.sample_period, mlog, engine,
                        sgc::write_no_status<calculator_type>);

  // oracle
  refcpu::IsingConfiguration oc(c.rows, c.cols, 1);
  for (int k = 0; k < 5; ++k) oc.occupation[(k * 7919) % (c.rows * c.cols)] = -1;
  refcpu::CompletionCheckParams q;
  if (c.max_count > 0) q.cutoff_params.max_count = c.max_count;
  if (c.min_sample > 0) q.cutoff_params.min_sample = c.min_sample;
  q.check_begin = 100;
  q.check_period = 10;
  if (c.prec > 0) {
    refcpu::RequestedPrecision rp;
    rp.abs_convergence_is_required = true;
    rp.abs_precision = c.prec;
    q.requested_precision[refcpu::SamplerComponent{"potential_energy", 0}] = rp;
    q.requested_precision[refcpu::SamplerComponent{"param_composition", 0}] = rp;
  }
  auto oeng = std::make_shared<std::mt19937_64>(c.seed);
  auto o = refcpu::sgc_basic_run(oc, 0.1, c.use_nlist, c.T, c.mu, oeng, q, c.sample_period);

  std::string tag = std::to_string(c.rows) + "x" + std::to_string(c.cols) + " T=" +
                    std::to_string((int)c.T) + " sp=" + std::to_string(c.sample_period);
  CountType ns = get_n_samples(data->samplers);
  EXPECT(ns == refcpu::get_n_samples(o.samplers), tag << " n_samples " << ns);
  for (auto name : {"formation_energy", "param_composition", "potential_energy"}) {
    VectorXd a = data->samplers.at(name)->component(0);
    auto b = o.samplers.at(name)->component(0);
    bool ok = a.size() == (Index)b.size();
    for (Index i = 0; ok && i < a.size(); ++i) ok = same_bits(a(i), b[i]);
    EXPECT(ok, tag << " sampled " << name);
  }
  if (c.host_function) {
    VectorXd m = data->samplers.at("magnetisation")->component(0);
    bool ok = m.size() == (Index)o.M.size();
    for (Index i = 0; ok && i < m.size(); ++i) ok = m(i) == (double)o.M[i];
    EXPECT(ok, tag << " host sampling function");
  }
  EXPECT(data->n_pass == o.n_pass, tag << " n_pass " << data->n_pass << " vs " << o.n_pass);
  EXPECT(data->n_accept == o.n_accept && data->n_reject == o.n_reject, tag << " counts");
  EXPECT(state.configuration.occupation() == o.final_config.occupation, tag << " final config");
  EXPECT(*engine == *oeng, tag << " engine state");
  auto const &ra = data->completion_check.results();
  auto const &rb = o.results;
  EXPECT(ra.is_complete == rb.is_complete && ra.n_samples == rb.n_samples &&
             ra.equilibration_check_results.all_equilibrated ==
                 rb.equilibration_check_results.all_equilibrated &&
             ra.equilibration_check_results.N_samples_for_all_to_equilibrate ==
                 rb.equilibration_check_results.N_samples_for_all_to_equilibrate &&
             ra.convergence_check_results.all_converged ==
                 rb.convergence_check_results.all_converged &&
             data->completion_check.n_checks() == o.n_checks,
         tag << " completion results");
  for (auto const &p : rb.convergence_check_results.individual_results) {
    auto it = ra.convergence_check_results.individual_results.find(
        SamplerComponent(p.first.sampler_name, p.first.component_index, ""));
    EXPECT(it != ra.convergence_check_results.individual_results.end() &&
               same_bits(it->second.stats.mean, p.second.stats.mean) &&
               same_bits(it->second.stats.calculated_precision, p.second.stats.calculated_precision),
           tag << " stats " << p.first.sampler_name);
  }
  std::cout << "gpu: " << tag << " n_samples=" << ns << " n_pass=" << data->n_pass
            << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// ising_cpp::complete_semigrand_canonical: the reference test's run set-up
// (Ising_complete_semigrand_canonical_test.cpp:270-345: 25x25 all +1, T 2000 K,
// mu 0, J 0.1, the row/column fold, min_sample 100, checks from 100 every 10,
// abs precision 0.001 on param_composition and potential_energy, the JSON
// configuration sampler, a status writer held in a std::function), with a
// seeded engine, vs the oracle's loop bit for bit.
static void gpu_complete_case(bool use_nlist, uint64_t seed, int sample_period) {
  namespace csgc = CASM::monte::ising_cpp::complete_semigrand_canonical;
  typedef csgc::SemiGrandCanonicalEventGenerator<std::mt19937_64> gen_type;
  typedef csgc::SemiGrandCanonicalCalculator<system_type, gen_type> calc_type;
  auto system = std::make_shared<system_type>(ising_cpp::IsingFormationEnergy(0.1, 1, use_nlist),
                                              ising_cpp::IsingParamComposition());
  auto calc = std::make_shared<calc_type>(system);
  StateSamplingFunctionMap functions;
  for (auto f : {csgc::make_parametric_composition_f(calc), csgc::make_formation_energy_f(calc),
                 csgc::make_potential_energy_f(calc)})
    functions.emplace(f.name, f);
  jsonStateSamplingFunctionMap json_functions;
  {
    auto jf = csgc::make_configuration_json_f(calc);
    json_functions.emplace(jf.name, jf);
  }
  CompletionCheckParams<BasicStatistics> ccp;
  ccp.cutoff_params.min_sample = 100;
  ccp.log_spacing = false;
  ccp.check_begin = 100;
  ccp.check_period = 10;
  converge(functions, ccp).set_precision("param_composition", 0.001)
      .set_precision("potential_energy", 0.001);
  ValueMap cond;
  cond.scalar_values["temperature"] = 2000.0;
  cond.vector_values["exchange_potential"] = VectorXd{0.0};
  ising_cpp::IsingState state(ising_cpp::IsingConfiguration({25, 25}, 1), cond);
  auto engine = std::make_shared<std::mt19937_64>(seed);
  std::optional<MethodLog> mlog = MethodLog();
  mlog->log_frequency = 10.0;
  int n_status = 0;
  std::function<void(calc_type const &, MethodLog &)> status = [&](calc_type const &c,
                                                                    MethodLog &ml) {
    ++n_status;
    csgc::write_no_status(c, ml);
  };
  calc->run(state, functions, json_functions, ccp, gen_type(), sample_period, mlog, engine,
            status);
  auto const &data = *calc->data;

  refcpu::IsingConfiguration oc(25, 25, 1);
  refcpu::CompletionCheckParams q;
  q.cutoff_params.min_sample = 100;
  q.check_begin = 100;
  q.check_period = 10;
  refcpu::RequestedPrecision rp;
  rp.abs_convergence_is_required = true;
  rp.abs_precision = 0.001;
  q.requested_precision[refcpu::SamplerComponent{"potential_energy", 0}] = rp;
  q.requested_precision[refcpu::SamplerComponent{"param_composition", 0}] = rp;
  auto oeng = std::make_shared<std::mt19937_64>(seed);
  auto o = refcpu::sgc_basic_run(oc, 0.1, use_nlist, 2000.0, 0.0, oeng, q, sample_period);

  const std::string tag = std::string("complete_semigrand_canonical ") +
                          (use_nlist ? "nlist" : "fold") + " seed " + std::to_string(seed) +
                          " sp " + std::to_string(sample_period);
  const CountType ns = get_n_samples(data.samplers);
  EXPECT(ns == refcpu::get_n_samples(o.samplers), tag << ": n_samples " << ns);
  for (auto name : {"formation_energy", "param_composition", "potential_energy"}) {
    VectorXd a = data.samplers.at(name)->component(0);
    auto b = o.samplers.at(name)->component(0);
    bool ok = a.size() == (Index)b.size();
    for (Index i = 0; ok && i < a.size(); ++i) ok = same_bits(a(i), b[i]);
    EXPECT(ok, tag << ": sampled " << name);
  }
  EXPECT((CountType)data.json_samplers.at("configuration").values.size() == ns,
         tag << ": one configuration per sample");
  EXPECT(data.n_steps_per_pass == 625, tag << ": n_steps_per_pass = n_sites");
  EXPECT(data.n_pass == o.n_pass && data.n_accept == o.n_accept && data.n_reject == o.n_reject,
         tag << ": counts");
  EXPECT(state.configuration.occupation() == o.final_config.occupation, tag << ": configuration");
  EXPECT(*engine == *oeng, tag << ": engine state");
  auto const &ra = data.completion_check.results();
  EXPECT(ra.is_complete && ra.is_complete == o.results.is_complete &&
             ra.convergence_check_results.all_converged ==
                 o.results.convergence_check_results.all_converged &&
             data.completion_check.n_checks() == o.n_checks,
         tag << ": completion results");
  EXPECT(n_status >= 1, tag << ": status written at the end");
  jsonParser json;
  csgc::to_json(data, json);
  EXPECT(json.is_obj() && json.contains("completion_check_results") &&
             json["n_steps_per_pass"].get<long>() == 625,
         tag << ": data to_json");
  std::cout << "gpu: " << tag << " n_samples=" << ns << " n_pass=" << data.n_pass
            << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// occupation_metropolis with a fixture sampling by step (period 37 steps on a
// 12x10 lattice, so samples fall inside passes; a max_count of 3053 steps cuts
// the run inside a pass) beside a fixture sampling every pass: the samples,
// counts, final configuration and engine equal the oracle's step-by-step loop.
static void gpu_occupation_metropolis_by_step() {
  const int R = 12, C = 10, N = R * C;
  const long S = 3053;
  const double T = 2300.0, mu = 0.05;
  const uint64_t seed = 9;
  typedef ising_cpp::IsingConfiguration config_type;
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  cond.vector_values["exchange_potential"] = VectorXd{mu};
  State<config_type> state(config_type({R, C}, 1), cond);
  IsingSemiGrandCanonicalPotential potential(0.1, true);
  potential.set(&state);
  StateSamplingFunction pe("potential_energy", "running potential energy", {}, [&state]() {
    return VectorXd{state.properties.scalar_values.at("potential_energy")};
  });
  pe.device_observable = MCG_OBS_POTENTIAL_ENERGY;
  StateSamplingFunction mag("magnetisation", "sum of spins", {}, [&state]() {
    double s = 0;
    for (int v : state.configuration.occupation()) s += v;
    return VectorXd{s};
  });
  StateSamplingFunctionMap fns;
  fns.emplace(pe.name, pe);
  fns.emplace(mag.name, mag);
  SamplingParams by_step;
  by_step.sample_mode = SAMPLE_MODE::BY_STEP;
  by_step.sampler_names = {"potential_energy", "magnetisation"};
  by_step.begin = 5.0;
  by_step.period = 37.0;
  CompletionCheckParams<BasicStatistics> cc_step;
  cc_step.cutoff_params.max_count = S;
  SamplingParams by_pass;
  by_pass.sampler_names = {"potential_energy"};
  CompletionCheckParams<BasicStatistics> cc_pass;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("steps", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), by_step,
                        cc_step, std::vector<std::string>{});
  fixtures.emplace_back("passes", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), by_pass,
                        cc_pass, std::vector<std::string>{});
  auto engine = std::make_shared<std::mt19937_64>(seed);
  rm_type rm(engine, fixtures, true);
  RandomNumberGenerator<std::mt19937_64> rng(engine);
  occupation_metropolis(state, potential, rng, rm);

  auto e1 = std::make_shared<std::mt19937_64>(seed);
  auto o = refcpu::sgc_step_run(refcpu::IsingConfiguration(R, C, 1), 0.1, true, T, mu, e1, S);
  const std::string tag = "occupation_metropolis by step";
  auto const &a = rm.sampling_fixtures[0]->results();
  VectorXd ape = a.samplers.at("potential_energy")->component(0);
  VectorXd am = a.samplers.at("magnetisation")->component(0);
  const long n_a = (S - 5) / 37 + 1;  // counts 5, 42, ..., <= S
  bool ok = ape.size() == n_a && am.size() == n_a && (long)a.sample_count.size() == n_a;
  for (long k = 0; ok && k < n_a; ++k) {
    const long g = 5 + 37 * k;
    ok = same_bits(ape(k), o.pe[g]) && am(k) == (double)o.M[g] && a.sample_count[k] == g;
  }
  EXPECT(ok, tag << ": step samples (" << ape.size() << " of " << n_a << ")");
  auto const &b = rm.sampling_fixtures[1]->results();
  VectorXd bpe = b.samplers.at("potential_energy")->component(0);
  const long n_b = S / N;  // passes 1..25
  ok = bpe.size() == n_b;
  for (long p = 1; ok && p <= n_b; ++p) ok = same_bits(bpe(p - 1), o.pe[p * N]);
  EXPECT(ok, tag << ": pass samples (" << bpe.size() << " of " << n_b << ")");
  auto const &ca = rm.sampling_fixtures[0]->counter();
  auto const &cb = rm.sampling_fixtures[1]->counter();
  EXPECT(ca.count == S && ca.pass == S / N && ca.step == S % N && cb.count == S / N,
         tag << ": counters (" << ca.count << ", " << ca.pass << ", " << ca.step << ")");
  EXPECT(state.configuration.occupation() == o.final_config.occupation, tag << ": configuration");
  EXPECT(*engine == *e1, tag << ": engine state");
  EXPECT(a.n_accept + a.n_reject == S, tag << ": attempts");
  std::cout << "gpu: " << tag << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// Stochastic sampling periods: a by-step and a by-pass fixture draw their next
// sample counts from the run's engine between steps of the walk, so the walk's
// stream and the sample counts both depend on where those draws fall.
static void gpu_occupation_metropolis_stochastic() {
  const int R = 10, C = 12, N = R * C;
  const long S = 2503;
  const double T = 2600.0, mu = -0.05;
  const uint64_t seed = 21;
  typedef ising_cpp::IsingConfiguration config_type;
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  cond.vector_values["exchange_potential"] = VectorXd{mu};
  State<config_type> state(config_type({R, C}, 1), cond);
  IsingSemiGrandCanonicalPotential potential(0.1, false);
  potential.set(&state);
  StateSamplingFunction pe("potential_energy", "running potential energy", {}, [&state]() {
    return VectorXd{state.properties.scalar_values.at("potential_energy")};
  });
  pe.device_observable = MCG_OBS_POTENTIAL_ENERGY;
  StateSamplingFunctionMap fns;
  fns.emplace(pe.name, pe);
  SamplingParams by_step;
  by_step.sample_mode = SAMPLE_MODE::BY_STEP;
  by_step.sampler_names = {"potential_energy"};
  by_step.begin = 3.0;
  by_step.period = 25.0;
  by_step.stochastic_sample_period = true;
  CompletionCheckParams<BasicStatistics> cc_step;
  cc_step.cutoff_params.max_count = S;
  SamplingParams by_pass;
  by_pass.sampler_names = {"potential_energy"};
  by_pass.period = 2.0;
  by_pass.stochastic_sample_period = true;
  CompletionCheckParams<BasicStatistics> cc_pass;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("steps", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), by_step,
                        cc_step, std::vector<std::string>{});
  fixtures.emplace_back("passes", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), by_pass,
                        cc_pass, std::vector<std::string>{});
  auto engine = std::make_shared<std::mt19937_64>(seed);
  rm_type rm(engine, fixtures, true);
  RandomNumberGenerator<std::mt19937_64> rng(engine);
  occupation_metropolis(state, potential, rng, rm);

  // the oracle's walk with the two fixtures' draws after each step, in
  // fixture order (SamplingFixture.hh:529-534, SamplingParams.hh:242-272)
  auto e1 = std::make_shared<std::mt19937_64>(seed);
  refcpu::RandomNumberGenerator orng(e1);
  long next_a = 3, next_b = 1;
  std::vector<long> counts_a, counts_b;
  auto o = refcpu::sgc_step_run(
      refcpu::IsingConfiguration(R, C, 1), 0.1, false, T, mu, e1, S, [&](long g) {
        if (g == next_a) {
          counts_a.push_back(g);
          next_a = g + refcpu::stochastic_count_step(1.0 / 25.0, orng);
        }
        if (g % N == 0 && g / N == next_b) {
          counts_b.push_back(g / N);
          next_b = g / N + refcpu::stochastic_count_step(1.0 / 2.0, orng);
        }
      });
  const std::string tag = "occupation_metropolis stochastic periods";
  auto const &a = rm.sampling_fixtures[0]->results();
  VectorXd ape = a.samplers.at("potential_energy")->component(0);
  bool ok = ape.size() == (Index)counts_a.size() && a.sample_count.size() == counts_a.size() &&
            counts_a.size() > 50;
  for (size_t k = 0; ok && k < counts_a.size(); ++k)
    ok = a.sample_count[k] == counts_a[k] && same_bits(ape(k), o.pe[counts_a[k]]);
  EXPECT(ok, tag << ": step samples (" << ape.size() << " of " << counts_a.size() << ")");
  auto const &b = rm.sampling_fixtures[1]->results();
  VectorXd bpe = b.samplers.at("potential_energy")->component(0);
  ok = bpe.size() == (Index)counts_b.size() && b.sample_count.size() == counts_b.size() &&
       counts_b.size() > 3;
  for (size_t k = 0; ok && k < counts_b.size(); ++k)
    ok = b.sample_count[k] == counts_b[k] && same_bits(bpe(k), o.pe[counts_b[k] * N]);
  EXPECT(ok, tag << ": pass samples (" << bpe.size() << " of " << counts_b.size() << ")");
  EXPECT(state.configuration.occupation() == o.final_config.occupation, tag << ": configuration");
  EXPECT(*engine == *e1, tag << ": engine state");
  EXPECT(a.n_accept + a.n_reject == S, tag << ": attempts");
  std::cout << "gpu: " << tag << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// Results files of a run on the MI355X (jsonResultsIO as the fixture's
// results_io): run.0/observations.json holds the sampled series bit for bit as
// the oracle's walk gives it, and summary.json the conditions and the
// BasicStatistics of that series (the oracle's calculator on the oracle's
// series, same bits).
static void gpu_results_files() {
  const int R = 16, C = 16, N = R * C;
  const long P = 400;
  const double T = 1900.0, mu = 0.1;
  const uint64_t seed = 31;
  typedef ising_cpp::IsingConfiguration config_type;
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  cond.vector_values["exchange_potential"] = VectorXd{mu};
  State<config_type> state(config_type({R, C}, 1), cond);
  IsingSemiGrandCanonicalPotential potential(0.1, true);
  potential.set(&state);
  StateSamplingFunction pe("potential_energy", "running potential energy", {}, [&state]() {
    return VectorXd{state.properties.scalar_values.at("potential_energy")};
  });
  pe.device_observable = MCG_OBS_POTENTIAL_ENERGY;
  StateSamplingFunctionMap fns;
  fns.emplace(pe.name, pe);
  SamplingParams sp;
  sp.sampler_names = {"potential_energy"};
  CompletionCheckParams<BasicStatistics> cc;
  cc.cutoff_params.max_count = P;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  typedef Results<config_type, BasicStatistics> results_type;
  std::filesystem::path dir = std::filesystem::temp_directory_path() /
                              ("mcg_results_" + std::to_string(::getpid()));
  std::filesystem::remove_all(dir);
  std::filesystem::create_directories(dir);
  jsonResultsIO<results_type> io(dir, false, true);
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("passes", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                        std::vector<std::string>{}, io.writer());
  auto engine = std::make_shared<std::mt19937_64>(seed);
  rm_type rm(engine, fixtures, true);
  RandomNumberGenerator<std::mt19937_64> rng(engine);
  occupation_metropolis(state, potential, rng, rm);

  auto e1 = std::make_shared<std::mt19937_64>(seed);
  auto o = refcpu::sgc_step_run(refcpu::IsingConfiguration(R, C, 1), 0.1, true, T, mu, e1, P * N);
  std::vector<double> series;
  for (long p = 1; p <= P; ++p) series.push_back(o.pe[p * N]);
  auto st = refcpu::basic_statistics(series, 0.95);

  const std::string tag = "results files from a GPU run";
  jsonParser obs = jsonParser::read(dir / "run.0" / "observations.json");
  auto vals = obs["potential_energy"]["value"].get<std::vector<double>>();
  auto counts = obs["count"].get<std::vector<long>>();
  bool ok = vals.size() == series.size() && counts.size() == series.size();
  for (size_t k = 0; ok && k < series.size(); ++k)
    ok = same_bits(vals[k], series[k]) && counts[k] == (long)k + 1;
  EXPECT(ok, tag << ": observations.json (" << vals.size() << " samples)");
  jsonParser sum = jsonParser::read(dir / "summary.json");
  auto means = sum["statistics"]["potential_energy"]["value"]["mean"].get<std::vector<double>>();
  auto precs = sum["statistics"]["potential_energy"]["value"]["calculated_precision"]
                   .get<std::vector<double>>();
  EXPECT(means.size() == 1 && precs.size() == 1 && same_bits(means[0], st.mean) &&
             same_bits(precs[0], st.calculated_precision),
         tag << ": summary.json statistics");
  auto temps = sum["conditions"]["temperature"]["value"].get<std::vector<double>>();
  EXPECT(temps.size() == 1 && temps[0] == T, tag << ": summary.json conditions");
  std::filesystem::remove_all(dir);
  std::cout << "gpu: " << tag << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// RunManager-driven loop: two fixtures, the FP64 running potential energy and
// a configuration-reading function, vs the oracle.
static void gpu_occupation_metropolis_tests() {
  const int R = 20, C = 20;
  const long P = 300;
  const double T = 2500.0, mu = 0.1;
  const uint64_t seed = 5;
  typedef ising_cpp::IsingConfiguration config_type;
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  cond.vector_values["exchange_potential"] = VectorXd{mu};
  State<config_type> state(config_type({R, C}, 1), cond);
  IsingSemiGrandCanonicalPotential potential(0.1, true);
  potential.set(&state);

  StateSamplingFunction pe("potential_energy", "running potential energy", {}, [&state]() {
    return VectorXd{state.properties.scalar_values.at("potential_energy")};
  });
  pe.device_observable = MCG_OBS_POTENTIAL_ENERGY;  // reads no configuration
  StateSamplingFunction mag("magnetisation", "sum of spins", {}, [&state]() {
    double s = 0;
    for (int v : state.configuration.occupation()) s += v;
    return VectorXd{s};
  });
  StateSamplingFunctionMap fns;
  fns.emplace(pe.name, pe);
  fns.emplace(mag.name, mag);

  SamplingParams sp1;
  sp1.sampler_names = {"potential_energy"};
  CompletionCheckParams<BasicStatistics> cc1;
  cc1.cutoff_params.max_count = P;
  SamplingParams sp2;
  sp2.sampler_names = {"magnetisation"};
  sp2.begin = 0.0;
  sp2.period = 7.0;
  CompletionCheckParams<BasicStatistics> cc2;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("thermo", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp1, cc1,
                        std::vector<std::string>{});
  fixtures.emplace_back("sparse", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp2, cc2,
                        std::vector<std::string>{});
  auto engine = std::make_shared<std::mt19937_64>(seed);
  rm_type rm(engine, fixtures, true);
  RandomNumberGenerator<std::mt19937_64> rng(engine);
  occupation_metropolis(state, potential, rng, rm);

  // oracle: running energy per pass and integer observables per pass
  auto e1 = std::make_shared<std::mt19937_64>(seed);
  double pe0 = 0.0;
  auto run_e = refcpu::sgc_running_energy_run(refcpu::IsingConfiguration(R, C, 1), 0.1, true, T,
                                              mu, e1, P, &pe0);
  auto e2 = std::make_shared<std::mt19937_64>(seed);
  refcpu::CompletionCheckParams q;
  q.cutoff_params.max_count = P;
  auto o = refcpu::sgc_basic_run(refcpu::IsingConfiguration(R, C, 1), 0.1, true, T, mu, e2, q, 1);

  auto const &thermo = rm.sampling_fixtures[0]->results();
  VectorXd got = thermo.samplers.at("potential_energy")->component(0);
  bool ok = got.size() == P;
  for (long i = 0; ok && i < P; ++i) ok = same_bits(got(i), run_e[i]);
  EXPECT(ok, "occupation_metropolis running energy (" << got.size() << " samples)");
  auto const &sparse = rm.sampling_fixtures[1]->results();
  VectorXd m = sparse.samplers.at("magnetisation")->component(0);
  ok = m.size() == (P / 7) + 1 && m(0) == (double)(R * C);
  for (Index k = 1; ok && k < m.size(); ++k) ok = m(k) == (double)o.M[7 * k - 1];
  EXPECT(ok, "occupation_metropolis second fixture (" << m.size() << " samples)");
  EXPECT(sparse.sample_count.size() == (size_t)m.size() && sparse.sample_count[1] == 7,
         "sample counts");
  EXPECT(*engine == *e1, "occupation_metropolis engine state");
  EXPECT(state.configuration.occupation() == o.final_config.occupation, "final configuration");
  EXPECT(thermo.n_accept == o.n_accept && thermo.n_reject == o.n_reject, "acceptance counts");
  EXPECT(thermo.completion_check_results.is_complete, "thermo fixture complete");
  std::cout << "gpu: occupation_metropolis RunManager 20x20, 2 fixtures"
            << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// ---------------------------------------------------------------------------
// Canonical pair swaps: host event machinery (CPU) and the device run (GPU)
// ---------------------------------------------------------------------------
static std::vector<int> random_occupants(int n, int n_one, uint64_t seed) {
  std::vector<int> occ(n, 0);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::mt19937_64 e(seed);
  std::shuffle(idx.begin(), idx.end(), e);
  for (int i = 0; i < n_one; ++i) occ[idx[i]] = 1;
  return occ;
}

// The invariants the reference's OccLocation test checks after every
// proposal and apply (tests/unit/monte/OccLocation_test.cpp:16-66).
static bool occ_location_consistent(OccLocation const &ol, std::vector<int> const &occ) {
  for (Index id = 0; id < ol.mol_size(); ++id)
    if (ol.mol(id).id != id) return false;
  for (Index l = 0; l < (Index)occ.size(); ++l) {
    Mol const &m = ol.mol(ol.l_to_mol_id(l));
    if (m.l != l || m.species_index != occ[l]) return false;
    if (ol.mol_id(ol.candidate_list().index(m.asym, m.species_index), m.loc) != m.id) return false;
  }
  return true;
}

static void cpu_canonical_tests() {
  const int R = 9, C = 7, N = R * C;
  ising_cpp::BinaryLatticeEvents ev(N);
  EXPECT(ev.candidate_list.size() == 2 && ev.canonical_swaps.size() == 1 &&
             ev.canonical_swaps[0] == OccSwap(OccCandidate(0, 0), OccCandidate(0, 1)),
         "binary candidate list and canonical swap");
  EXPECT(ev.semigrand_canonical_swaps.size() == 2, "semi-grand canonical swaps (both orders)");
  std::vector<int> occ = random_occupants(N, 25, 3);
  OccLocation ol(ev.convert, ev.candidate_list);
  ol.initialize(occ);
  EXPECT(occ_location_consistent(ol, occ), "OccLocation::initialize invariants");
  // the oracle's restatement of the same proposal and apply, same stream
  refcpu::BinaryOccLocation ro;
  ro.initialize(occ);
  std::vector<int> rocc = occ;
  auto eng = std::make_shared<std::mt19937_64>(77);
  std::mt19937_64 reng(77);
  RandomNumberGenerator<std::mt19937_64> rng(eng);
  OccEvent e;
  refcpu::CanonicalEvent re;
  bool same = true, inv = true;
  for (int k = 0; k < 20000; ++k) {
    propose_canonical_event(e, ol, ev.canonical_swaps, rng);
    refcpu::propose_canonical_event(re, ro, reng);
    same = same && e.linear_site_index[0] == re.l[0] && e.linear_site_index[1] == re.l[1] &&
           e.new_occ[0] == re.new_occ[0] && e.new_occ[1] == re.new_occ[1];
    if (k % 3 != 2) {  // apply two of every three proposals
      ol.apply(e, occ);
      for (int i = 0; i < 2; ++i) ro.apply_transform(re.mol_id[i], re.from[i], re.to[i], rocc);
    }
    if (k % 500 == 0) inv = inv && occ_location_consistent(ol, occ);
  }
  EXPECT(same, "propose_canonical_event matches the oracle's stream use");
  EXPECT(inv && occ_location_consistent(ol, occ), "OccLocation::apply invariants");
  EXPECT(occ == rocc && ol.locations()[0] == ro.loc[0] && ol.locations()[1] == ro.loc[1],
         "lists after 13334 applied swaps");
  EXPECT(*eng == reng, "engine state after 20000 proposals");
  // semi-grand canonical proposal: one site of the chosen type changes species
  propose_semigrand_canonical_event(e, ol, ev.semigrand_canonical_swaps, rng);
  EXPECT(e.linear_site_index.size() == 1 && e.new_occ[0] == 1 - occ[e.linear_site_index[0]],
         "semi-grand canonical proposal flips one site");
  std::vector<int> all_zero(N, 0);
  ol.initialize(all_zero);
  bool threw = false;
  try {
    propose_canonical_event(e, ol, ev.canonical_swaps, rng);
  } catch (std::runtime_error const &ex) {
    threw = std::string(ex.what()).find("No events possible") != std::string::npos;
  }
  EXPECT(threw, "no canonical event possible at x = 0");
  std::cout << "cpu: canonical event machinery" << (g_fail ? " (failures so far)" : " ok")
            << std::endl;
}

static void gpu_canonical_tests() {
  const int R = 18, C = 14, N = R * C;
  const long P = 45;
  const double T = 1700.0, J = 0.1;
  const uint64_t seed = 31;
  typedef ising_cpp::IsingConfiguration config_type;
  std::vector<int> occ0 = random_occupants(N, 101, 9);
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  State<config_type> state(cfg, cond);
  ising_cpp::IsingOccupantPotential potential(J, true);
  potential.set(&state);
  ising_cpp::BinaryLatticeEvents ev(N);
  OccLocation ol(ev.convert, ev.candidate_list);
  ol.initialize(state.configuration.occupation());

  StateSamplingFunction pe("potential_energy", "running potential energy", {}, [&state]() {
    return VectorXd{state.properties.scalar_values.at("potential_energy")};
  });
  pe.device_observable = MCG_OBS_POTENTIAL_ENERGY;
  StateSamplingFunctionMap fns;
  fns.emplace(pe.name, pe);
  SamplingParams sp;
  sp.sampler_names = {"potential_energy"};
  CompletionCheckParams<BasicStatistics> cc;
  cc.cutoff_params.max_count = P;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("thermo", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                        std::vector<std::string>{});
  auto engine = std::make_shared<std::mt19937_64>(seed);
  rm_type rm(engine, fixtures, true);
  typedef RandomNumberGenerator<std::mt19937_64> generator_type;
  generator_type rng(engine);
  occupation_metropolis(state, ol, potential, ev.canonical_swaps,
                        propose_canonical_event<generator_type>, rng, rm);

  std::mt19937_64 oe(seed);
  auto o = refcpu::canonical_run(R, C, occ0, J, T, oe, P);
  auto const &res = rm.sampling_fixtures[0]->results();
  VectorXd got = res.samplers.at("potential_energy")->component(0);
  bool ok = got.size() == P;
  for (long i = 0; ok && i < P; ++i) ok = same_bits(got(i), o.e_run[i]);
  EXPECT(ok, "canonical running energy (" << got.size() << " samples)");
  EXPECT(state.configuration.occupation() == o.final_occ, "canonical final configuration");
  EXPECT(ol.locations()[0] == o.ol.loc[0] && ol.locations()[1] == o.ol.loc[1],
         "canonical OccLocation lists");
  EXPECT(occ_location_consistent(ol, state.configuration.occupation()),
         "canonical OccLocation invariants after the run");
  EXPECT(*engine == oe, "canonical engine state");
  EXPECT(res.n_accept == o.n_accept && res.n_reject == o.n_reject, "canonical acceptance counts");
  std::cout << "gpu: canonical occupation_metropolis 18x14" << (g_fail ? " (failures so far)" : " ok")
            << std::endl;
}

static bool close_rel(double a, double b, double rel) {
  return std::fabs(a - b) <= rel * std::max(std::fabs(a), std::fabs(b));
}

// nfold() on the MI355X (NFOLD mode) vs refcpu::nfold_run.  Two ways to stop:
//  * max_count = P: the check at count P comes before that pass's selection,
//    so the samples are counts 0..P-1 and the engine has drawn exactly the
//    events of P passes (oracle: P passes, no boundary sample but the first);
//  * convergence: the check right after the sample at count c stops the run
//    after that pass's first event (oracle: c passes sampled every pass, then
//    its pending event applied).
static void gpu_nfold_case(bool converge_stop) {
  const int R = 10, C = 12, N = R * C;
  const long P = 40;
  const double T = 1800.0, J = 0.1, mu = 0.05;
  const uint64_t seed = converge_stop ? 52 : 51;
  typedef ising_cpp::IsingConfiguration config_type;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  typedef NfoldData<config_type, BasicStatistics, std::mt19937_64> data_type;
  std::vector<int> occ0 = random_occupants(N, 57, 13);
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  cond.vector_values["exchange_potential"] = VectorXd{mu};
  State<config_type> state(cfg, cond);
  ising_cpp::BinaryLatticeEvents ev(N);
  OccLocation ol(ev.convert, ev.candidate_list);
  ol.initialize(state.configuration.occupation());
  data_type nfold_data;
  nfold_data.n_events_possible = N;

  StateSamplingFunctionMap fns;
  StateSamplingFunction m("spin_sum", "sum of spins", {}, [&state]() {
    double s = 0;
    for (int v : state.configuration.occupation()) s += 2 * v - 1;
    return VectorXd{s};
  });
  StateSamplingFunction ear("expected_acceptance_rate", "total rate / N", {}, [&nfold_data]() {
    return VectorXd{nfold_data.expected_acceptance_rate};
  });
  fns.emplace(m.name, m);
  fns.emplace(ear.name, ear);
  SamplingParams sp;
  sp.sampler_names = {"spin_sum", "expected_acceptance_rate"};
  sp.do_sample_time = true;
  // the cutoff case samples from count 0 (the initial state, weighted by the
  // first event's increment), the convergence case from count 1
  sp.begin = converge_stop ? 1.0 : 0.0;
  const long first = converge_stop ? 1 : 0;
  CompletionCheckParams<BasicStatistics> cc;
  if (converge_stop) {
    cc.cutoff_params.max_count = 20000;
    cc.check_begin = 50;
    cc.check_period = 10;
    converge(fns, cc).set_precision("spin_sum", 4.0);
  } else {
    cc.cutoff_params.max_count = P;
  }
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("kinetics", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                        std::vector<std::string>{});
  auto rm_engine = std::make_shared<std::mt19937_64>(seed + 1000);
  rm_type rm(rm_engine, fixtures, true);
  auto engine = std::make_shared<std::mt19937_64>(seed);
  ising_cpp::IsingNfoldEventSelector<std::mt19937_64> selector(J, engine);
  OccEvent flip;
  flip.occ_transform.resize(1);
  auto get_event_f = [&](Index l) -> OccEvent const & {
    Mol const &mol = ol.mol(ol.l_to_mol_id(l));
    flip.linear_site_index = {l};
    flip.new_occ = {1 - state.configuration.occ(l)};
    flip.occ_transform[0] = OccTransform{l, mol.id, mol.asym, mol.species_index,
                                         1 - mol.species_index};
    return flip;
  };
  nfold<Index>(state, ol, nfold_data, selector, get_event_f, rm);

  auto const &f = *rm.sampling_fixtures[0];
  auto const &res = f.results();
  const long n_samp = res.samplers.at("spin_sum")->n_samples();
  const long c_stop = f.counter().count;
  const std::string tag = converge_stop ? "nfold convergence stop" : "nfold count cutoff";
  refcpu::IsingConfiguration oc(R, C, -1);
  for (int l = 0; l < N; ++l) oc.occupation[l] = occ0[l] ? 1 : -1;
  // the samples: counts first..first+n_samp-1
  std::mt19937_64 e1(seed);
  auto o1 = refcpu::nfold_run(oc, J, T, mu, e1, first + n_samp - 1, 1, 0.0, first == 0);
  bool ok = (long)o1.M.size() == n_samp && (long)res.sample_weight.n_samples() == n_samp &&
            (long)res.sample_time.size() == n_samp;
  VectorXd got_m = res.samplers.at("spin_sum")->component(0);
  VectorXd got_r = res.samplers.at("expected_acceptance_rate")->component(0);
  VectorXd got_w = res.sample_weight.component(0);
  for (long i = 0; ok && i < n_samp; ++i)
    ok = got_m(i) == (double)o1.M[i] && close_rel(got_w(i), o1.weight[i], 1e-12) &&
         close_rel(got_r(i), o1.rate[i] / N, 1e-13) &&
         close_rel(res.sample_time[i], o1.time[i], 1e-12) && res.sample_count[i] == first + i;
  EXPECT(ok, tag << ": samples (" << n_samp << ")");
  if (converge_stop) {
    EXPECT(c_stop == n_samp && f.counter().step == 1 && c_stop < 20000,
           tag << ": stops one event into the pass of its last sample (count " << c_stop
               << ", step " << f.counter().step << ")");
    std::vector<int> want(o1.final_occ.size());
    for (int l = 0; l < N; ++l) want[l] = o1.final_occ[l] > 0 ? 1 : 0;
    want[o1.pending_site] = 1 - want[o1.pending_site];
    EXPECT(state.configuration.occupation() == want, tag << ": final configuration");
    EXPECT(*engine == e1, tag << ": engine state");
    EXPECT(res.n_accept == (BigCountType)c_stop * N + 1 && res.n_reject == 0,
           tag << ": event count " << res.n_accept);
    EXPECT(close_rel(f.counter().time, o1.final_time + o1.pending_dt, 1e-12), tag << ": time");
    // the stopping check saw the time after that event (SamplingFixture.hh:271-279)
    auto const &ct = f.completion_check().results().time;
    EXPECT(ct.has_value() && close_rel(*ct, o1.final_time + o1.pending_dt, 1e-12),
           tag << ": completion check time");
  } else {
    EXPECT(n_samp == P && c_stop == P && f.counter().step == 0, tag << ": stops at count P");
    std::mt19937_64 e2(seed);
    auto o2 = refcpu::nfold_run(oc, J, T, mu, e2, P, P + 1, 0.0, true);
    std::vector<int> want(N);
    for (int l = 0; l < N; ++l) want[l] = o2.final_occ[l] > 0 ? 1 : 0;
    EXPECT(state.configuration.occupation() == want, tag << ": final configuration");
    EXPECT(*engine == e2, tag << ": engine state (no selection at count P)");
    EXPECT(res.n_accept == (BigCountType)P * N && res.n_reject == 0, tag << ": event count");
    EXPECT(close_rel(f.counter().time, o2.final_time, 1e-12), tag << ": time");
    auto const &ct = f.completion_check().results().time;
    EXPECT(ct.has_value() && close_rel(*ct, o2.final_time, 1e-12),
           tag << ": completion check time");
  }
  EXPECT(*rm_engine == std::mt19937_64(seed + 1000), tag << ": run manager engine untouched");
  EXPECT(occ_location_consistent(ol, state.configuration.occupation()),
         tag << ": OccLocation invariants after the run");
  std::cout << "gpu: " << tag << " 10x12 (" << n_samp << " samples)"
            << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// One site (N = 1), a fixture sampling every 3rd pass from count 0 with
// max_sample = 1: the reference samples count 0 (selecting the event), applies
// that event (which ends pass 0), and its check at count 1 meets max_sample.
// The run must stop there: count 1, one event, the engine after one selection.
static void gpu_nfold_one_site_max_sample() {
  const double T = 1500.0, J = 0.1, mu = 0.02;
  const uint64_t seed = 71;
  typedef ising_cpp::IsingConfiguration config_type;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  config_type cfg({1, 1}, 0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  cond.vector_values["exchange_potential"] = VectorXd{mu};
  State<config_type> state(cfg, cond);
  ising_cpp::BinaryLatticeEvents ev(1);
  OccLocation ol(ev.convert, ev.candidate_list);
  ol.initialize(state.configuration.occupation());
  NfoldData<config_type, BasicStatistics, std::mt19937_64> nfold_data;
  nfold_data.n_events_possible = 1;
  StateSamplingFunctionMap fns;
  StateSamplingFunction m("spin_sum", "sum of spins", {}, [&state]() {
    return VectorXd{2.0 * state.configuration.occ(0) - 1.0};
  });
  fns.emplace(m.name, m);
  SamplingParams sp;
  sp.sampler_names = {"spin_sum"};
  sp.begin = 0.0;
  sp.period = 3.0;
  sp.do_sample_time = true;
  CompletionCheckParams<BasicStatistics> cc;
  cc.cutoff_params.max_sample = 1;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("one", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                        std::vector<std::string>{});
  rm_type rm(std::make_shared<std::mt19937_64>(seed + 1), fixtures, true);
  auto engine = std::make_shared<std::mt19937_64>(seed);
  ising_cpp::IsingNfoldEventSelector<std::mt19937_64> selector(J, engine);
  OccEvent flip;
  flip.occ_transform.resize(1);
  auto get_event_f = [&](Index l) -> OccEvent const & {
    Mol const &mol = ol.mol(ol.l_to_mol_id(l));
    flip.linear_site_index = {l};
    flip.new_occ = {1 - state.configuration.occ(l)};
    flip.occ_transform[0] = OccTransform{l, mol.id, mol.asym, mol.species_index,
                                         1 - mol.species_index};
    return flip;
  };
  nfold<Index>(state, ol, nfold_data, selector, get_event_f, rm);
  auto const &f = *rm.sampling_fixtures[0];
  auto const &res = f.results();
  // oracle: the count-0 sample, then one pass (its one event) and no sample at count 1
  refcpu::IsingConfiguration oc(1, 1, -1);
  std::mt19937_64 e1(seed);
  auto o = refcpu::nfold_run(oc, J, T, mu, e1, 1, 3, 0.0, true);
  const std::string tag = "nfold one site max_sample 1";
  EXPECT(res.samplers.at("spin_sum")->n_samples() == 1 && o.M.size() == 1 &&
             res.samplers.at("spin_sum")->component(0)(0) == (double)o.M[0],
         tag << ": the count-0 sample");
  EXPECT(f.counter().count == 1 && f.counter().step == 0,
         tag << ": stops at count 1 (count " << f.counter().count << ")");
  EXPECT(res.n_accept == 1, tag << ": one event (" << res.n_accept << ")");
  EXPECT(state.configuration.occ(0) == (o.final_occ[0] > 0 ? 1 : 0), tag << ": configuration");
  EXPECT(*engine == e1, tag << ": engine after one selection");
  EXPECT(close_rel(f.counter().time, o.final_time, 1e-12), tag << ": time");
  std::cout << "gpu: " << tag << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// Two fixtures sampling every pass and every 4th pass from count 0, both cut
// off at count 30: each holds the oracle's samples at its own counts (the
// device records only the passes where some fixture is due).
// nfold() with a fixture sampling by time (period tau from tau, weight 1,
// the state before the event whose time reaches each sample time) beside a
// fixture sampling by pass that ends the run at count P: both against the
// oracle's time-grid and by-pass runs of the same chain (sampling draws
// nothing from the engine, so both runs follow one trajectory).
static void gpu_nfold_by_time() {
  const int R = 10, C = 12, N = R * C;
  const long P = 30;
  const double T = 1600.0, J = 0.1, mu = -0.03, tau = 0.37;
  const uint64_t seed = 81;
  typedef ising_cpp::IsingConfiguration config_type;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<int> occ0 = random_occupants(N, 43, 17);
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  cond.vector_values["exchange_potential"] = VectorXd{mu};
  State<config_type> state(cfg, cond);
  ising_cpp::BinaryLatticeEvents ev(N);
  OccLocation ol(ev.convert, ev.candidate_list);
  ol.initialize(state.configuration.occupation());
  NfoldData<config_type, BasicStatistics, std::mt19937_64> nfold_data;
  nfold_data.n_events_possible = N;
  StateSamplingFunctionMap fns;
  StateSamplingFunction m("spin_sum", "sum of spins", {}, [&state]() {
    double s = 0;
    for (int v : state.configuration.occupation()) s += 2 * v - 1;
    return VectorXd{s};
  });
  StateSamplingFunction ear("expected_acceptance_rate", "total rate / N", {}, [&nfold_data]() {
    return VectorXd{nfold_data.expected_acceptance_rate};
  });
  fns.emplace(m.name, m);
  fns.emplace(ear.name, ear);
  SamplingParams sp_time;
  sp_time.sampler_names = {"spin_sum", "expected_acceptance_rate"};
  sp_time.sample_mode = SAMPLE_MODE::BY_TIME;
  sp_time.begin = tau;
  sp_time.period = tau;
  sp_time.do_sample_time = true;
  SamplingParams sp_pass;
  sp_pass.sampler_names = {"spin_sum"};
  sp_pass.begin = 1.0;
  sp_pass.do_sample_time = true;
  CompletionCheckParams<BasicStatistics> cc_pass;
  cc_pass.cutoff_params.max_count = P;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("by_time", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp_time,
                        CompletionCheckParams<BasicStatistics>(), std::vector<std::string>{});
  fixtures.emplace_back("by_pass", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp_pass,
                        cc_pass, std::vector<std::string>{});
  rm_type rm(std::make_shared<std::mt19937_64>(seed + 1000), fixtures, true);
  auto engine = std::make_shared<std::mt19937_64>(seed);
  ising_cpp::IsingNfoldEventSelector<std::mt19937_64> selector(J, engine);
  OccEvent flip;
  flip.occ_transform.resize(1);
  auto get_event_f = [&](Index l) -> OccEvent const & {
    Mol const &mol = ol.mol(ol.l_to_mol_id(l));
    flip.linear_site_index = {l};
    flip.new_occ = {1 - state.configuration.occ(l)};
    flip.occ_transform[0] = OccTransform{l, mol.id, mol.asym, mol.species_index,
                                         1 - mol.species_index};
    return flip;
  };
  nfold<Index>(state, ol, nfold_data, selector, get_event_f, rm);

  const std::string tag = "nfold by time";
  refcpu::IsingConfiguration oc(R, C, -1);
  for (int l = 0; l < N; ++l) oc.occupation[l] = occ0[l] ? 1 : -1;
  std::mt19937_64 eg(seed), ep(seed);
  auto og = refcpu::nfold_run(oc, J, T, mu, eg, P, 1, tau);
  auto op = refcpu::nfold_run(oc, J, T, mu, ep, P - 1, 1, 0.0, false);
  // the grid records expanded: one sample per grid time, weight 1
  std::vector<double> want_m, want_t;
  for (size_t r = 0; r < og.M.size(); ++r)
    for (long k = 0; k < (long)og.weight[r]; ++k) {
      want_m.push_back((double)og.M[r]);
      want_t.push_back(og.time[r] + k * tau);
    }
  auto const &rt = rm.sampling_fixtures[0]->results();
  VectorXd got_m = rt.samplers.at("spin_sum")->component(0);
  VectorXd got_w = rt.sample_weight.component(0);
  bool ok = (size_t)got_m.size() == want_m.size() && rt.sample_time.size() == want_m.size() &&
            (size_t)got_w.size() == want_m.size() && !want_m.empty();
  for (size_t i = 0; ok && i < want_m.size(); ++i)
    ok = got_m(i) == want_m[i] && got_w(i) == 1.0 &&
         close_rel(rt.sample_time[i], (double)(i + 1) * tau, 1e-12) &&
         close_rel(rt.sample_time[i], want_t[i], 1e-9);
  EXPECT(ok, tag << ": samples by time (" << got_m.size() << " of " << want_m.size() << ")");
  VectorXd got_r = rt.samplers.at("expected_acceptance_rate")->component(0);
  bool rate_ok = got_r.size() == (Index)want_m.size();
  for (Index i = 0; rate_ok && i < got_r.size(); ++i) rate_ok = got_r(i) > 0.0 && got_r(i) <= 1.0;
  EXPECT(rate_ok, tag << ": expected acceptance rates in (0, 1]");
  // by pass: counts 1..P-1 (the run ends at count P before sampling it)
  auto const &rp = rm.sampling_fixtures[1]->results();
  VectorXd pm = rp.samplers.at("spin_sum")->component(0);
  VectorXd pw = rp.sample_weight.component(0);
  bool okp = pm.size() == P - 1 && (long)op.M.size() == P - 1;
  for (long i = 0; okp && i < P - 1; ++i)
    okp = pm(i) == (double)op.M[i] && close_rel(pw(i), op.weight[i], 1e-12) &&
          rp.sample_count[i] == i + 1;
  EXPECT(okp, tag << ": samples by pass (" << pm.size() << ")");
  std::vector<int> want(N);
  for (int l = 0; l < N; ++l) want[l] = og.final_occ[l] > 0 ? 1 : 0;
  EXPECT(state.configuration.occupation() == want, tag << ": final configuration");
  EXPECT(*engine == eg, tag << ": engine state");
  EXPECT(rp.n_accept == (BigCountType)P * N, tag << ": event count " << rp.n_accept);
  EXPECT(close_rel(rm.sampling_fixtures[1]->counter().time, og.final_time, 1e-12),
         tag << ": time");
  EXPECT(occ_location_consistent(ol, state.configuration.occupation()),
         tag << ": OccLocation invariants after the run");
  std::cout << "gpu: " << tag << " 10x12 (" << got_m.size() << " time samples)"
            << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

static void gpu_nfold_two_fixtures() {
  const int R = 6, C = 9, N = R * C;
  const long P = 30;
  const double T = 2200.0, J = 0.1, mu = -0.1;
  const uint64_t seed = 61;
  typedef ising_cpp::IsingConfiguration config_type;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<int> occ0 = random_occupants(N, 20, 17);
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  cond.vector_values["exchange_potential"] = VectorXd{mu};
  State<config_type> state(cfg, cond);
  ising_cpp::BinaryLatticeEvents ev(N);
  OccLocation ol(ev.convert, ev.candidate_list);
  ol.initialize(state.configuration.occupation());
  NfoldData<config_type, BasicStatistics, std::mt19937_64> nfold_data;
  nfold_data.n_events_possible = N;
  StateSamplingFunctionMap fns;
  StateSamplingFunction m("spin_sum", "sum of spins", {}, [&state]() {
    double v = 0;
    for (int o : state.configuration.occupation()) v += 2 * o - 1;
    return VectorXd{v};
  });
  fns.emplace(m.name, m);
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  for (double period : {1.0, 4.0}) {
    SamplingParams sp;
    sp.sampler_names = {"spin_sum"};
    sp.begin = 0.0;
    sp.period = period;
    CompletionCheckParams<BasicStatistics> cc;
    cc.cutoff_params.max_count = P;
    fixtures.emplace_back(period == 1.0 ? "every" : "sparse", fns, jsonStateSamplingFunctionMap(),
                          ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                          std::vector<std::string>{});
  }
  auto engine = std::make_shared<std::mt19937_64>(seed);
  rm_type rm(engine, fixtures, false);
  ising_cpp::IsingNfoldEventSelector<std::mt19937_64> selector(J, engine);
  OccEvent e;
  nfold<Index>(state, ol, nfold_data, selector, [&](Index) -> OccEvent const & { return e; }, rm);

  refcpu::IsingConfiguration oc(R, C, -1);
  for (int l = 0; l < N; ++l) oc.occupation[l] = occ0[l] ? 1 : -1;
  std::mt19937_64 e1(seed);
  auto o1 = refcpu::nfold_run(oc, J, T, mu, e1, P - 1, 1, 0.0, true);
  for (int fi = 0; fi < 2; ++fi) {
    auto const &res = rm.sampling_fixtures[fi]->results();
    const long step = fi == 0 ? 1 : 4, want_n = fi == 0 ? P : (P - 1) / 4 + 1;
    VectorXd got = res.samplers.at("spin_sum")->component(0);
    VectorXd w = res.sample_weight.component(0);
    bool ok = got.size() == want_n && w.size() == want_n;
    for (long i = 0; ok && i < want_n; ++i)
      ok = res.sample_count[i] == i * step && got(i) == (double)o1.M[i * step] &&
           close_rel(w(i), o1.weight[i * step], 1e-12);
    EXPECT(ok, "nfold two fixtures: fixture " << fi << " samples (" << got.size() << ")");
  }
  std::mt19937_64 e2(seed);
  auto o2 = refcpu::nfold_run(oc, J, T, mu, e2, P, P + 1, 0.0, true);
  std::vector<int> want(N);
  for (int l = 0; l < N; ++l) want[l] = o2.final_occ[l] > 0 ? 1 : 0;
  EXPECT(state.configuration.occupation() == want, "nfold two fixtures: final configuration");
  EXPECT(*engine == e2, "nfold two fixtures: engine state");
  std::cout << "gpu: nfold two fixtures 6x9" << (g_fail ? " (failures so far)" : " ok")
            << std::endl;
}

// nfold's argument checks run before any device call
static void cpu_nfold_argument_tests() {
  typedef ising_cpp::IsingConfiguration config_type;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  const int N = 16;
  auto attempt = [&](SAMPLE_MODE mode, bool time_cutoff, int bad_occ) {
    config_type cfg({4, 4}, 0);
    if (bad_occ) cfg.set_occ(3, bad_occ);
    ValueMap cond;
    cond.scalar_values["temperature"] = 1000.0;
    cond.vector_values["exchange_potential"] = VectorXd{0.0};
    State<config_type> state(cfg, cond);
    ising_cpp::BinaryLatticeEvents ev(N);
    OccLocation ol(ev.convert, ev.candidate_list);
    ol.initialize(std::vector<int>(N, 0));
    SamplingParams sp;
    sp.sample_mode = mode;
    sp.do_sample_time = time_cutoff;
    CompletionCheckParams<BasicStatistics> cc;
    cc.cutoff_params.max_count = 10;
    if (time_cutoff) cc.cutoff_params.max_time = 1.0;
    std::vector<rm_type::sampling_fixture_params_type> fixtures;
    fixtures.emplace_back("f", StateSamplingFunctionMap(), jsonStateSamplingFunctionMap(),
                          ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                          std::vector<std::string>{});
    auto engine = std::make_shared<std::mt19937_64>(1);
    rm_type rm(engine, fixtures, true);
    NfoldData<config_type, BasicStatistics, std::mt19937_64> data;
    ising_cpp::IsingNfoldEventSelector<std::mt19937_64> selector(0.1, engine);
    OccEvent e;
    try {
      nfold<Index>(state, ol, data, selector, [&](Index) -> OccEvent const & { return e; }, rm);
    } catch (std::runtime_error const &err) {
      return std::string(err.what());
    }
    return std::string();
  };
  EXPECT(attempt(SAMPLE_MODE::BY_STEP, false, 0).find("BY_PASS") != std::string::npos,
         "nfold rejects BY_STEP fixtures");
  EXPECT(attempt(SAMPLE_MODE::BY_PASS, true, 0).find("max_time") != std::string::npos,
         "nfold rejects simulated-time cutoffs");
  EXPECT(attempt(SAMPLE_MODE::BY_PASS, false, -1).find("0/1") != std::string::npos,
         "nfold rejects spin configurations");
  std::cout << "cpu: nfold argument checks" << (g_fail ? " (failures so far)" : " ok")
            << std::endl;
}

// ---------------------------------------------------------------------------
// Several asymmetric units: units (A, B), (B, A) and (A) on a square lattice
// (the unit of each site drawn at random)
// ---------------------------------------------------------------------------
static std::vector<Index> random_units(int n, uint64_t seed) {
  std::mt19937_64 e(seed);
  std::vector<Index> ua(n);
  for (auto &b : ua) b = (Index)(e() % 3);
  return ua;
}

static std::vector<int> unit_occupants(std::vector<Index> const &ua, uint64_t seed) {
  std::mt19937_64 e(seed);
  std::vector<int> occ(ua.size());
  for (size_t l = 0; l < ua.size(); ++l) occ[l] = ua[l] == 2 ? 0 : (int)(e() % 2);
  return occ;
}

static const std::vector<std::vector<std::string>> kThreeUnits = {{"A", "B"}, {"B", "A"}, {"A"}};

// OccLocation_test.cpp:16-66 for several units: fixed sites map to mol_size(),
// every Mol in the list of its unit and species at its loc
static bool occ_location_consistent_units(OccLocation const &ol, std::vector<int> const &occ) {
  auto const &cv = ol.convert();
  for (Index id = 0; id < ol.mol_size(); ++id)
    if (ol.mol(id).id != id) return false;
  for (Index l = 0; l < (Index)occ.size(); ++l) {
    const Index b = cv.l_to_asym(l);
    if (cv.occ_size(b) < 2) {
      if (ol.l_to_mol_id(l) != ol.mol_size()) return false;
      continue;
    }
    Mol const &m = ol.mol(ol.l_to_mol_id(l));
    if (m.l != l || m.asym != b || m.species_index != cv.species_index(b, occ[l])) return false;
    if (ol.mol_id(ol.candidate_list().index(m.asym, m.species_index), m.loc) != m.id) return false;
  }
  return true;
}

static refcpu::OccSystem oracle_system(std::vector<Index> const &ua) {
  return refcpu::OccSystem(std::vector<int>(ua.begin(), ua.end()), {{0, 1}, {1, 0}, {0}}, 2);
}

static std::vector<std::pair<int, int>> swap_indices(OccCandidateList const &cl,
                                                     std::vector<OccSwap> const &swaps) {
  std::vector<std::pair<int, int>> out;
  for (auto const &s : swaps) out.push_back({(int)cl.index(s.cand_a), (int)cl.index(s.cand_b)});
  return out;
}

static void cpu_occupant_tests() {
  const int R = 8, C = 9, N = R * C;
  auto ua = random_units(N, 5);
  Conversions cv(ua, kThreeUnits);
  OccCandidateList cl(cv);
  EXPECT(cv.species_size() == 2 && cv.occ_index(1, 0) == 1 && !cv.species_allowed(2, 1),
         "Conversions with three units");
  EXPECT(cl.size() == 4 && cl[2] == OccCandidate(1, 1) && cl[3] == OccCandidate(1, 0),
         "candidate list in unit and occupant order");
  auto sgc = make_semigrand_canonical_swaps(cv, cl);
  auto can = make_canonical_swaps(cv, cl);
  refcpu::OccSystem rs = oracle_system(ua);
  EXPECT(swap_indices(cl, sgc) == refcpu::occ_semigrand_swaps(rs), "semi-grand swap list");
  EXPECT(swap_indices(cl, can) == refcpu::occ_canonical_swaps(rs), "canonical swap list");
  const double J = 0.1;
  const std::vector<double> mu = {0.02, -0.05};
  for (int canonical = 0; canonical < 2; ++canonical) {
    auto const &swaps = canonical ? can : sgc;
    auto idx = swap_indices(cl, swaps);
    std::vector<int> occ = unit_occupants(ua, 6 + canonical);
    ising_cpp::IsingConfiguration cfg({R, C}, 0);
    cfg.set_occupation(occ);
    State<ising_cpp::IsingConfiguration> state(cfg, ValueMap());
    ising_cpp::IsingOccupantPotential pot(J, cv, mu, true);
    pot.set(&state);
    OccLocation ol(cv, cl);
    ol.initialize(state.configuration.occupation());
    refcpu::OccLocationG ro;
    ro.initialize(rs, occ);
    std::vector<int> rocc = occ;
    refcpu::IsingConfiguration shape(R, C, 1);
    EXPECT(same_bits(pot.per_supercell(),
                     refcpu::occ_potential_per_supercell(shape, rs, rocc, J, mu, true)),
           "potential per supercell");
    auto eng = std::make_shared<std::mt19937_64>(77 + canonical);
    std::mt19937_64 reng(77 + canonical);
    RandomNumberGenerator<std::mt19937_64> rng(eng);
    OccEvent e;
    refcpu::OccEventG re;
    bool same = true, delta = true, inv = true;
    for (int k = 0; k < 20000; ++k) {
      if (canonical)
        propose_canonical_event(e, ol, swaps, rng);
      else
        propose_semigrand_canonical_event(e, ol, swaps, rng);
      refcpu::occ_propose(re, ro, idx, canonical != 0, reng);
      same = same && (int)e.linear_site_index.size() == re.n;
      for (int i = 0; same && i < re.n; ++i)
        same = e.linear_site_index[i] == re.l[i] && e.new_occ[i] == re.new_occ[i];
      delta = delta && same_bits(pot.occ_delta_per_supercell(e.linear_site_index, e.new_occ),
                                 refcpu::occ_potential_delta(shape, rocc, J, mu, re));
      if (k % 3 != 2) {
        ol.apply(e, state.configuration.occupation());
        for (int i = 0; i < re.n; ++i) ro.apply_transform(re.mol_id[i], re.from[i], re.to[i], rocc);
      }
      if (k % 500 == 0) inv = inv && occ_location_consistent_units(ol, state.configuration.occupation());
    }
    std::string tag = canonical ? "canonical" : "semi-grand";
    EXPECT(same, tag << " proposals match the oracle's stream use");
    EXPECT(delta, tag << " potential deltas equal the oracle's bit for bit");
    EXPECT(inv && occ_location_consistent_units(ol, state.configuration.occupation()),
           tag << " OccLocation invariants");
    bool lists = state.configuration.occupation() == rocc;
    for (Index c = 0; lists && c < cl.size(); ++c)
      lists = ol.locations()[c] == ro.loc[c];
    EXPECT(lists, tag << " occupation and lists after the applied events");
    EXPECT(*eng == reng, tag << " engine state after 20000 proposals");
  }
  std::cout << "cpu: occupant events over three units" << (g_fail ? " (failures so far)" : " ok")
            << std::endl;
}

// occupation_metropolis through the reference's generic signature with
// propose_semigrand_canonical_event or propose_canonical_event over three
// units (OCCUPANT mode) vs refcpu::occ_run
static void gpu_occupant_case(bool canonical, double T, uint64_t seed) {
  const int R = 14, C = 12, N = R * C;
  const long P = 30;
  const double J = 0.1;
  const std::vector<double> mu = {0.0, 0.04};
  typedef ising_cpp::IsingConfiguration config_type;
  auto ua = random_units(N, seed);
  Conversions cv(ua, kThreeUnits);
  OccCandidateList cl(cv);
  auto swaps = canonical ? make_canonical_swaps(cv, cl) : make_semigrand_canonical_swaps(cv, cl);
  std::vector<int> occ0 = unit_occupants(ua, seed + 1);
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  State<config_type> state(cfg, cond);
  ising_cpp::IsingOccupantPotential potential(J, cv, mu, true);
  potential.set(&state);
  OccLocation ol(cv, cl);
  ol.initialize(state.configuration.occupation());
  const Index n_mut = ol.mol_size();

  StateSamplingFunction pe("potential_energy", "running potential energy", {}, [&state]() {
    return VectorXd{state.properties.scalar_values.at("potential_energy")};
  });
  pe.device_observable = MCG_OBS_POTENTIAL_ENERGY;
  StateSamplingFunctionMap fns;
  fns.emplace(pe.name, pe);
  SamplingParams sp;
  sp.sampler_names = {"potential_energy"};
  CompletionCheckParams<BasicStatistics> cc;
  cc.cutoff_params.max_count = P;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("thermo", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                        std::vector<std::string>{});
  auto engine = std::make_shared<std::mt19937_64>(seed);
  rm_type rm(engine, fixtures, true);
  typedef RandomNumberGenerator<std::mt19937_64> generator_type;
  generator_type rng(engine);
  if (canonical)
    occupation_metropolis(state, ol, potential, swaps, propose_canonical_event<generator_type>, rng,
                          rm);
  else
    occupation_metropolis(state, ol, potential, swaps,
                          propose_semigrand_canonical_event<generator_type>, rng, rm);

  refcpu::OccSystem rs = oracle_system(ua);
  std::mt19937_64 oe(seed);
  auto o = refcpu::occ_run(R, C, rs, occ0, J, mu, swap_indices(cl, swaps), canonical, T, N, oe,
                           P * n_mut);
  std::string tag = std::string(canonical ? "canonical" : "semi-grand") + " three units";
  auto const &res = rm.sampling_fixtures[0]->results();
  VectorXd got = res.samplers.at("potential_energy")->component(0);
  bool ok = got.size() == P && o.error.empty();
  for (long i = 0; ok && i < P; ++i) ok = same_bits(got(i), o.e_run[i]);
  EXPECT(ok, tag << ": running energy (" << got.size() << " samples)");
  EXPECT(state.configuration.occupation() == o.final_occ, tag << ": final configuration");
  bool lists = true;
  for (Index c = 0; lists && c < cl.size(); ++c) lists = ol.locations()[c] == o.ol.loc[c];
  EXPECT(lists, tag << ": OccLocation lists");
  EXPECT(occ_location_consistent_units(ol, state.configuration.occupation()),
         tag << ": OccLocation invariants after the run");
  EXPECT(*engine == oe, tag << ": engine state");
  EXPECT(res.n_accept == o.n_accept && res.n_reject == o.n_reject, tag << ": acceptance counts");
  std::cout << "gpu: occupation_metropolis, " << tag << (g_fail ? " (failures so far)" : " ok")
            << std::endl;
}

// a canonical run whose only swap type runs out of events: the loop stops
// there (occupation_metropolis.hh:118-127) with the steps made so far
static void gpu_occupant_no_events() {
  const int R = 6, C = 6, N = R * C;
  const double T = 50000.0;
  typedef ising_cpp::IsingConfiguration config_type;
  auto ua = random_units(N, 18);
  Conversions cv(ua, kThreeUnits);
  OccCandidateList cl(cv);
  std::vector<OccSwap> swaps = {OccSwap(OccCandidate(0, 0), OccCandidate(1, 1))};
  std::vector<int> occ0 = unit_occupants(ua, 19);
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  State<config_type> state(cfg, cond);
  ising_cpp::IsingOccupantPotential potential(0.0, true);
  potential.set(&state);
  OccLocation ol(cv, cl);
  ol.initialize(state.configuration.occupation());
  StateSamplingFunction pe("potential_energy", "running potential energy", {}, [&state]() {
    return VectorXd{state.properties.scalar_values.at("potential_energy")};
  });
  pe.device_observable = MCG_OBS_POTENTIAL_ENERGY;
  StateSamplingFunctionMap fns;
  fns.emplace(pe.name, pe);
  SamplingParams sp;
  sp.sampler_names = {"potential_energy"};
  CompletionCheckParams<BasicStatistics> cc;
  cc.cutoff_params.max_count = 100;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("thermo", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                        std::vector<std::string>{});
  auto engine = std::make_shared<std::mt19937_64>(20);
  rm_type rm(engine, fixtures, true);
  typedef RandomNumberGenerator<std::mt19937_64> generator_type;
  generator_type rng(engine);
  occupation_metropolis(state, ol, potential, swaps, propose_canonical_event<generator_type>, rng,
                        rm);
  refcpu::OccSystem rs = oracle_system(ua);
  std::mt19937_64 oe(20);
  auto o = refcpu::occ_run(R, C, rs, occ0, 0.0, {0.0, 0.0}, swap_indices(cl, swaps), true, T, N,
                           oe, 100 * ol.mol_size());
  EXPECT(o.error == "Error in choose_canonical_swap: No events possible.", "oracle stops");
  auto const &res = rm.sampling_fixtures[0]->results();
  EXPECT(res.n_accept == o.n_accept && res.n_reject == o.n_reject,
         "no-events stop: steps made (" << res.n_accept << " vs " << o.n_accept << ")");
  EXPECT(rm.sampling_fixtures[0]->counter().step + rm.sampling_fixtures[0]->counter().pass *
                                                       ol.mol_size() == o.n_steps,
         "no-events stop: the fixture's count");
  EXPECT(state.configuration.occupation() == o.final_occ, "no-events stop: configuration");
  EXPECT(*engine == oe, "no-events stop: engine");
  std::cout << "gpu: occupation_metropolis stops when no events are possible"
            << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// kinetic_monte_carlo() on the MI355X (KMC mode) vs refcpu::kmc_run: a fixture
// sampling by pass (completing the run at its max_count) and one sampling by
// time, with sampling functions that read the configuration and KMCData
static void gpu_kmc_case(uint64_t seed) {
  const int R = 10, C = 9, N = R * C;
  const long P = 12;
  const double J = 0.05, T = 900.0, tau = 0.75;
  typedef ising_cpp::IsingConfiguration config_type;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<int> occ0(N, 0);
  {
    std::mt19937_64 e(seed);
    for (auto &o : occ0) o = (e() % 100) < 35 ? 1 : 0;
  }
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  State<config_type> state(cfg, cond);
  Conversions cv(R, C, {"Va", "A"});
  OccCandidateList cl(cv);
  OccLocation ol(cv, cl, true);
  ol.initialize(state.configuration.occupation());
  const Index n_atoms = ol.atom_size();
  auto engine = std::make_shared<std::mt19937_64>(seed);
  ising_cpp::LatticeGasKmcEventSelector selector(J, T, engine, &state, &ol);
  KMCData<config_type, BasicStatistics, std::mt19937_64> kd;
  MatrixXd R0 = ol.atom_positions_cart();

  StateSamplingFunctionMap fns;
  fns.emplace("S", StateSamplingFunction("S", "bond sum of the atoms", {}, [&]() {
                auto const &c = state.configuration;
                double s = 0.0;
                for (Index l = 0; l < N; ++l)
                  if (c.occ(l))
                    s += c.occ(ising_cpp::lattice_gas_neighbour(l, 0, R, C)) +
                         c.occ(ising_cpp::lattice_gas_neighbour(l, 2, R, C));
                return VectorXd{s};
              }));
  fns.emplace("sum_sq", StateSamplingFunction("sum_sq", "sum of squared displacements", {}, [&]() {
                double s = 0.0;
                for (Index a = 0; a < kd.atom_positions_cart.cols(); ++a)
                  for (int k = 0; k < 3; ++k) {
                    const double d = kd.atom_positions_cart(k, a) - R0(k, a);
                    s += d * d;
                  }
                return VectorXd{s};
              }));
  fns.emplace("time", StateSamplingFunction("time", "KMC time", {}, [&]() { return VectorXd{kd.time}; }));
  SamplingParams by_pass;
  by_pass.sampler_names = {"S", "sum_sq", "time"};
  by_pass.begin = 0.0;
  CompletionCheckParams<BasicStatistics> cc_pass;
  cc_pass.cutoff_params.max_count = P;
  SamplingParams by_time = by_pass;
  by_time.sample_mode = SAMPLE_MODE::BY_TIME;
  by_time.period = tau;
  by_time.do_sample_time = true;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("by_pass", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), by_pass,
                        cc_pass, std::vector<std::string>{});
  fixtures.emplace_back("by_time", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), by_time,
                        CompletionCheckParams<BasicStatistics>(), std::vector<std::string>{});
  rm_type rm(engine, fixtures, true);
  kinetic_monte_carlo<ising_cpp::LatticeGasHop>(
      state, ol, kd, selector,
      [&](ising_cpp::LatticeGasHop h) -> OccEvent const & { return selector.get_event(h); }, rm);

  std::mt19937_64 e1(seed), e2(seed);
  auto o = refcpu::kmc_run(R, C, occ0, J, T, e1, P * N);
  auto og = refcpu::kmc_run(R, C, occ0, J, T, e2, P * N, tau);
  std::string tag = "kinetic_monte_carlo seed " + std::to_string(seed);
  auto const &rp = rm.sampling_fixtures[0]->results();
  VectorXd S = rp.samplers.at("S")->component(0), sq = rp.samplers.at("sum_sq")->component(0),
           tm = rp.samplers.at("time")->component(0);
  bool ok = S.size() == P;
  for (long c = 1; ok && c < P; ++c)
    ok = S(c) == (double)o.records[c - 1].S && sq(c) == (double)o.records[c - 1].sum_sq &&
         same_bits(tm(c), o.records[c - 1].time);
  ok = ok && sq(0) == 0.0 && tm(0) == 0.0;
  EXPECT(ok, tag << ": by-pass samples (" << S.size() << ")");
  auto const &rt = rm.sampling_fixtures[1]->results();
  VectorXd St = rt.samplers.at("S")->component(0), sqt = rt.samplers.at("sum_sq")->component(0),
           tt = rt.samplers.at("time")->component(0);
  bool okt = St.size() == (Index)og.records.size() && rt.sample_time.size() == og.records.size();
  for (Index k = 0; okt && k < St.size(); ++k)
    okt = St(k) == (double)og.records[k].S && sqt(k) == (double)og.records[k].sum_sq &&
          same_bits(tt(k), og.records[k].time) && same_bits(rt.sample_time[k], og.records[k].time);
  EXPECT(okt, tag << ": by-time samples (" << St.size() << " vs " << og.records.size() << ")");
  EXPECT(state.configuration.occupation() == o.final_occ, tag << ": final configuration");
  EXPECT(*engine == e1, tag << ": engine");
  bool atoms_ok = n_atoms == (Index)o.atoms.size();
  for (Index a = 0; atoms_ok && a < n_atoms; ++a) {
    auto const &at = ol.atom(a);
    atoms_ok = at.translation[0] - (long)R0(0, a) == o.atoms[a].dx &&
               at.translation[1] - (long)R0(1, a) == o.atoms[a].dy && at.n_jumps == o.atoms[a].n_jumps &&
               ol.mol(ol.l_to_mol_id(o.atoms[a].site)).component == std::vector<Index>{a};
  }
  EXPECT(atoms_ok, tag << ": atoms (translations, jumps, components)");
  bool lists_ok = ol.locations()[0] == std::vector<Index>(o.ol.loc[0].begin(), o.ol.loc[0].end()) &&
                  ol.locations()[1] == std::vector<Index>(o.ol.loc[1].begin(), o.ol.loc[1].end());
  EXPECT(lists_ok, tag << ": OccLocation lists");
  EXPECT(rp.n_accept == P * N, tag << ": events counted as accepted");
  std::cout << "gpu: " << tag << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

// the drop-in OccLocation's atom tracking against the oracle's, over the
// selector's own events applied on the host
// Selected-event statistics (SelectedEventFunctions.hh): every expected value
// below is worked out by hand from the binning rules (the first value places
// begin on the bin_width grid just below it, a value exactly on an edge goes
// to the bin above, smaller values prepend bins, at most max_size bins).
static void cpu_selected_event_tests() {
  {  // linear 1d histogram: growth both ways, then the max_size limit
    Histogram1D h(0.0, 1.0, false, 5);
    h.insert(2.5);
    EXPECT(h.begin() == 2.0 && h.count() == std::vector<double>({1.0}), "hist1d: first value");
    h.insert(0.3);
    EXPECT(h.begin() == 0.0 && h.count() == std::vector<double>({1.0, 0.0, 1.0}),
           "hist1d: bins prepended");
    h.insert(6.7);  // would need bin 6 of at most 5
    EXPECT(h.count() == std::vector<double>({1.0, 0.0, 1.0, 0.0, 0.0}) && h.max_size_exceeded() &&
               h.out_of_range_count() == 1.0,
           "hist1d: appended to max_size, the value out of range");
    h.insert(-2.2);  // no room below either
    EXPECT(h.begin() == 0.0 && h.out_of_range_count() == 2.0 && h.sum() == 4.0,
           "hist1d: below the full range");
    EXPECT(h.bin_coords() == std::vector<double>({0.0, 1.0, 2.0, 3.0, 4.0}) &&
               h.density() == std::vector<double>({0.25, 0.0, 0.25, 0.0, 0.0}),
           "hist1d: coordinates and density");
  }
  {  // log space: log10(10) = 1 lies on an edge and goes to the bin above it
    Histogram1D h(0.0, 0.5, true);
    h.insert(10.0);
    EXPECT(h.begin() == 0.5 && h.count() == std::vector<double>({0.0, 1.0}), "hist1d log: edge");
    h.insert_log_value(0.2, 2.0);
    EXPECT(h.begin() == 0.0 && h.count() == std::vector<double>({2.0, 0.0, 1.0}),
           "hist1d log: weighted log value");
    bool threw = false;
    try {
      Histogram1D(0.0, 1.0, false).insert_log_value(1.0);
    } catch (std::runtime_error const &) {
      threw = true;
    }
    EXPECT(threw, "hist1d: insert_log_value on a linear histogram throws");
  }
  {  // partitions: combined range, every partition aligned to it
    PartitionedHistogram1D p({"A", "B"}, 0.0, 1.0, false);
    p.insert(0, 1.5);
    p.insert(1, 4.2);
    p.insert(1, 4.7);
    auto const &c = p.combined_histogram();
    EXPECT(c.begin() == 1.0 && c.count() == std::vector<double>({1.0, 0.0, 0.0, 2.0}),
           "partitioned: combined histogram");
    auto const &hs = p.histograms();
    EXPECT(hs[0].begin() == 1.0 && hs[0].count() == std::vector<double>({1.0, 0.0, 0.0, 0.0}) &&
               hs[1].begin() == 1.0 && hs[1].count() == std::vector<double>({0.0, 0.0, 0.0, 2.0}),
           "partitioned: partitions span the combined range");
    jsonParser j;
    to_json(p, j);
    EXPECT(j["partition_names"].size() == 2 && j["partitions"].size() == 2 &&
               j["partitions"][1]["partition_fraction"].get<double>() == 2.0 / 3.0 &&
               j["combined"]["sum"].get<double>() == 3.0,
           "partitioned: JSON");
  }
  {  // discrete histograms: lexicographic keys, max_size, tolerance
    DiscreteVectorIntHistogram h({"a", "b"}, {2}, 2);
    h.insert(VectorXl{1, 2});
    h.insert(VectorXl{0, 5});
    h.insert(VectorXl{1, 2});
    h.insert(VectorXl{3, 3});
    EXPECT(h.size() == 2 && h.values()[0] == VectorXl({0, 5}) &&
               h.count() == std::vector<double>({1.0, 2.0}) && h.out_of_range_count() == 1.0 &&
               h.max_size_exceeded() && h.fraction() == std::vector<double>({0.25, 0.5}),
           "discrete int histogram");
    DiscreteVectorFloatHistogram f({"0"}, {}, 1e-5);
    f.insert(VectorXd{0.1});
    f.insert(VectorXd{0.1 + 1e-7});
    f.insert(VectorXd{0.2});
    EXPECT(f.size() == 2 && f.count() == std::vector<double>({2.0, 1.0}) && f.tol() == 1e-5,
           "discrete float histogram: values within tol share a bin");
    jsonParser j;
    to_json(h, j);
    EXPECT(j["values"].size() == 2 && j["values"][0][1].get<long>() == 5 &&
               j["out_of_range_count"].get<double>() == 1.0 && j["component_names"].size() == 2,
           "discrete histogram: JSON");
  }
  {  // hop correlations: every 2nd jump, at most 2 positions per atom
    CorrelationsData d;
    d.initialize(2, 2, 2, false);
    d.insert(0, 2, VectorXd{1.0, 0.0, 0.0}, 5, 1, 0.5);
    EXPECT(d.n_position_samples == std::vector<CountType>({1, 0}) && d.n_complete_samples == 0,
           "correlations: one atom sampled");
    d.insert(1, 1, VectorXd{0.0, 1.0, 0.0}, 6, 1, 0.7);  // odd jump count: not sampled
    d.insert(1, 2, VectorXd{0.0, 2.0, 0.0}, 7, 1, 0.9);
    EXPECT(d.n_position_samples == std::vector<CountType>({1, 1}) && d.n_complete_samples == 1 &&
               d.step(0, 1) == 7 && d.time(0, 1) == 0.9 && d.atom_positions_cart[0](1, 1) == 2.0,
           "correlations: both atoms, one complete sample");
    d.insert(0, 4, VectorXd{2.0, 0.0, 0.0}, 8, 2, 1.1);
    d.insert(0, 6, VectorXd{3.0, 0.0, 0.0}, 9, 2, 1.3);  // atom 0 full
    EXPECT(d.n_position_samples == std::vector<CountType>({2, 1}) && d.n_complete_samples == 1,
           "correlations: the full atom stops, the complete count waits for the other");
    jsonParser j;
    to_json(d, j);
    EXPECT(j["step"].size() == 1 && j["atom_positions_cart"].size() == 1 &&
               j["atom_positions_cart"][0].size() == 3,
           "correlations: JSON holds the complete samples");
  }
  {  // the collector: generic functions in order, histograms by name
    std::vector<std::string> trace;
    long type = 0;
    double rate = 1.0;
    SelectedEventFunctions fns;
    fns.insert(GenericSelectedEventFunction("second", "", false, [&]() { trace.push_back("second"); },
                                            []() { return true; }, 2));
    fns.insert(GenericSelectedEventFunction("first", "", false, [&]() { trace.push_back("first"); },
                                            []() { return true; }, 1));
    fns.insert(DiscreteVectorIntHistogramFunction("event_type", "", {}, std::nullopt, false,
                                                  [&]() { return VectorXl{type}; },
                                                  []() { return true; }, 100));
    fns.insert(PartitionedHistogramFunction<double>("rate", "", false, [&]() { return rate; },
                                                    {"all"}, []() { return 0; }, true, 0.0, 1.0,
                                                    100));
    SelectedEventFunctionParams params;
    params.evaluate("second", std::nullopt).evaluate("first", std::nullopt);
    params.collect("event_type", std::nullopt, std::nullopt, std::nullopt, std::nullopt, 3);
    params.collect("rate", std::nullopt, 0.5, std::nullopt, "log", std::nullopt);
    auto data = std::make_shared<SelectedEventData>();
    SelectedEventDataCollector collector(fns, params, data);
    for (long k = 0; k < 5; ++k) {
      type = k % 4;
      rate = std::pow(10.0, (double)k);
      collector.collect();
    }
    EXPECT(trace.size() == 10 && trace[0] == "first" && trace[1] == "second",
           "collector: generic functions by order");
    auto const &et = data->discrete_vector_int_histograms.at("event_type");
    EXPECT(et.max_size() == 3 && et.size() == 3 && et.out_of_range_count() == 1.0 &&
               et.count() == std::vector<double>({2.0, 1.0, 1.0}),
           "collector: discrete histogram with the params' max_size");
    auto const &rh = data->continuous_1d_histograms.at("rate").combined_histogram();
    EXPECT(rh.is_log() && rh.bin_width() == 0.5 && rh.sum() == 5.0 && rh.begin() == 0.0,
           "collector: log histogram with the params' bin width");
    bool threw = false;
    try {
      SelectedEventFunctionParams bad;
      bad.evaluate("missing", std::nullopt);
      SelectedEventDataCollector c2(fns, bad, data);
    } catch (std::runtime_error const &) {
      threw = true;
    }
    EXPECT(threw, "collector: unknown quantity throws");
  }
  {  // parameters: collect / do_not_collect and JSON round trip
    SelectedEventFunctionParams p;
    p.collect("event_type", std::nullopt, std::nullopt, std::nullopt, std::nullopt, std::nullopt);
    EXPECT(p.function_names == std::vector<std::string>({"event_type"}), "params: collect");
    p.do_not_collect("event_type");
    EXPECT(p.function_names.empty(), "params: do_not_collect");
    p.collect("event_type", std::nullopt, std::nullopt, std::nullopt, std::nullopt, std::nullopt);
    p.collect("dE_activated", std::nullopt, 1.0, 0.01, "linear", 10);
    p.correlations_data_params = CorrelationsDataParams{};
    p.correlations_data_params->jumps_per_position_sample = 10;
    jsonParser j;
    to_json(p, j);
    auto q = selected_event_function_params_from_json(j);
    EXPECT(q.function_names == p.function_names && q.bin_width == p.bin_width &&
               q.initial_begin == p.initial_begin && q.is_log == p.is_log &&
               q.max_size == p.max_size && q.correlations_data_params.has_value() &&
               q.correlations_data_params->jumps_per_position_sample == 10 &&
               j["spacing"]["dE_activated"].get<std::string>() == "linear",
           "params: JSON round trip");
    bool threw = false;
    try {
      p.collect("x", std::nullopt, std::nullopt, std::nullopt, "cubic", std::nullopt);
    } catch (std::runtime_error const &) {
      threw = true;
    }
    EXPECT(threw, "params: a spacing other than log or linear throws");
  }
  std::cout << "cpu: selected-event statistics" << (g_fail ? " (failures so far)" : " ok")
            << std::endl;
}

// kinetic_monte_carlo() with hop correlations recorded by the device (the
// selector's correlations_data_params): every 2nd jump, at most 3 positions
// per atom.  Each atom has min(3, jumps / 2) positions, the complete count is
// the least of them, an atom whose jumps are a multiple of 2 and that has
// fewer than 3 positions ends at its last stored position, positions and
// times only grow, and the run itself equals the oracle's.
static void gpu_kmc_hops() {
  const int R = 8, C = 7, N = R * C;
  const long P = 6;
  const double J = 0.05, T = 800.0;
  const uint64_t seed = 77;
  typedef ising_cpp::IsingConfiguration config_type;
  typedef RunManager<config_type, BasicStatistics, std::mt19937_64> rm_type;
  std::vector<int> occ0 = random_occupants(N, 20, 5);
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ0);
  ValueMap cond;
  cond.scalar_values["temperature"] = T;
  State<config_type> state(cfg, cond);
  Conversions cv(R, C, {"Va", "A"});
  OccCandidateList cl(cv);
  OccLocation ol(cv, cl, true);
  ol.initialize(state.configuration.occupation());
  const Index n_atoms = ol.atom_size();
  auto engine = std::make_shared<std::mt19937_64>(seed);
  ising_cpp::LatticeGasKmcEventSelector selector(J, T, engine, &state, &ol);
  selector.correlations_data_params = CorrelationsDataParams{2, 3, false, false};
  selector.selected_event_data = std::make_shared<SelectedEventData>();
  KMCData<config_type, BasicStatistics, std::mt19937_64> kd;
  StateSamplingFunctionMap fns;
  fns.emplace("time", StateSamplingFunction("time", "KMC time", {}, [&]() { return VectorXd{kd.time}; }));
  SamplingParams sp;
  sp.sampler_names = {"time"};
  CompletionCheckParams<BasicStatistics> cc;
  cc.cutoff_params.max_count = P;
  std::vector<rm_type::sampling_fixture_params_type> fixtures;
  fixtures.emplace_back("kmc", fns, jsonStateSamplingFunctionMap(),
                        ResultsAnalysisFunctionMap<config_type, BasicStatistics>(), sp, cc,
                        std::vector<std::string>{});
  rm_type rm(engine, fixtures, true);
  kinetic_monte_carlo<ising_cpp::LatticeGasHop>(
      state, ol, kd, selector,
      [&](ising_cpp::LatticeGasHop h) -> OccEvent const & { return selector.get_event(h); }, rm);
  const std::string tag = "kinetic_monte_carlo hop correlations";
  std::mt19937_64 e1(seed);
  auto o = refcpu::kmc_run(R, C, occ0, J, T, e1, P * N);
  EXPECT(*engine == e1, tag << ": engine state");
  auto const &cdo = selector.selected_event_data->correlations_data;
  EXPECT(cdo.has_value(), tag << ": data present");
  if (cdo.has_value()) {
    auto const &cd = *cdo;
    bool ok = (Index)cd.n_position_samples.size() == n_atoms;
    CountType least = 3;
    MatrixXd fin = ol.atom_positions_cart();
    for (Index a = 0; ok && a < n_atoms; ++a) {
      const Index jumps = ol.atom(a).n_jumps;
      const CountType want = std::min<CountType>(3, jumps / 2);
      ok = cd.n_position_samples[a] == want;
      least = std::min(least, want);
      for (CountType k = 1; ok && k < want; ++k)
        ok = cd.time(k, a) > cd.time(k - 1, a) &&
             cd.pass(k, a) * N + cd.step(k, a) > cd.pass(k - 1, a) * N + cd.step(k - 1, a);
      if (ok && want > 0 && want < 3 && jumps % 2 == 0)
        ok = cd.atom_positions_cart[want - 1](0, a) == fin(0, a) &&
             cd.atom_positions_cart[want - 1](1, a) == fin(1, a);
      if (ok && want > 0) ok = cd.time(want - 1, a) <= o.time;
    }
    EXPECT(ok && cd.n_complete_samples == least && cd.jumps_per_position_sample == 2,
           tag << ": records against the final atoms");
    jsonParser j;
    to_json(*selector.selected_event_data, j);
    EXPECT(j.contains("correlations_data") &&
               j["correlations_data"]["n_complete_samples"].get<long>() == (long)least,
           tag << ": JSON");
  }
  std::cout << "gpu: " << tag << " 8x7" << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

static void cpu_kmc_atom_tracking() {
  const int R = 6, C = 7, N = R * C;
  std::vector<int> occ(N, 0);
  for (int l = 0; l < N; l += 3) occ[l] = 1;
  typedef ising_cpp::IsingConfiguration config_type;
  config_type cfg({R, C}, 0);
  cfg.set_occupation(occ);
  ValueMap cond;
  cond.scalar_values["temperature"] = 1000.0;
  State<config_type> state(cfg, cond);
  Conversions cv(R, C, {"Va", "A"});
  OccCandidateList cl(cv);
  OccLocation ol(cv, cl, true);
  ol.initialize(state.configuration.occupation());
  ising_cpp::LatticeGasKmcEventSelector sel(0.05, 1000.0, std::make_shared<std::mt19937_64>(1),
                                            &state, &ol);
  std::mt19937_64 e(5);
  std::vector<std::array<long, 2>> disp(ol.atom_size(), {{0, 0}});
  std::vector<Index> where;
  for (Index l = 0; l < N; ++l)
    if (occ[l]) where.push_back(l);
  int applied = 0;
  for (int k = 0; k < 3000; ++k) {
    const Index from = where[e() % where.size()];
    const int dir = (int)(e() % 4);
    const Index to = ising_cpp::lattice_gas_neighbour(from, dir, R, C);
    if (state.configuration.occ(to)) continue;
    ol.apply(sel.get_event(ising_cpp::LatticeGasHop{from, dir}), state.configuration.occupation());
    const Index a = std::find(where.begin(), where.end(), from) - where.begin();
    where[a] = to;
    disp[a][0] += dir == 0 ? 1 : dir == 1 ? -1 : 0;
    disp[a][1] += dir == 2 ? 1 : dir == 3 ? -1 : 0;
    ++applied;
  }
  bool ok = applied > 500 && occ_location_consistent(ol, state.configuration.occupation());
  MatrixXd Rn = ol.atom_positions_cart();
  Index a0 = 0;
  for (Index l = 0; l < N; ++l)
    if (occ[l]) {
      ok = ok && Rn(0, a0) == (double)(l % R + disp[a0][0]) && Rn(1, a0) == (double)(l / R + disp[a0][1]) &&
           ol.mol(ol.l_to_mol_id(where[a0])).component == std::vector<Index>{a0};
      ++a0;
    }
  EXPECT(ok, "OccLocation atom tracking over " << applied << " hops");
  EXPECT(sel.total_rate() > 0.0, "lattice gas total rate");
  std::cout << "cpu: OccLocation atom tracking" << (g_fail ? " (failures so far)" : " ok") << std::endl;
}

static void gpu_tests() {
  gpu_complete_case(false, 42, 1);
  gpu_complete_case(true, 44, 2);
  gpu_occupation_metropolis_by_step();
  gpu_occupation_metropolis_stochastic();
  gpu_results_files();
  // the reference's SemiGrandCanonicalRun1 set-up, converged on two observables
  gpu_case({25, 25, true, 2000.0, 0.0, 42, 1, -1, 0.001, 100, false});
  gpu_case({25, 25, false, 2000.0, 0.0, 43, 1, -1, 0.001, 100, false});
  // cutoff-terminated, sample period 3, non-zero exchange potential
  gpu_case({16, 20, true, 3000.0, 0.2, 7, 3, 900, -1.0, -1, false});
  // a host-evaluated sampling function (lattice refreshed at every sample)
  gpu_case({12, 12, true, 2500.0, -0.1, 9, 2, 400, 0.002, 100, true});
  // tiny lattice: pending periodic checks span passes
  gpu_case({3, 3, true, 4000.0, 0.0, 11, 1, 3000, 1e-4, 500, false});
  gpu_occupation_metropolis_tests();
  gpu_canonical_tests();
  gpu_occupant_case(false, 1800.0, 41);
  gpu_occupant_case(true, 2200.0, 43);
  gpu_occupant_case(false, 400.0, 45);
  gpu_occupant_no_events();
  gpu_kmc_case(61);
  gpu_kmc_case(62);
  gpu_kmc_hops();
  gpu_nfold_case(false);
  gpu_nfold_case(true);
  gpu_nfold_two_fixtures();
  gpu_nfold_one_site_max_sample();
  gpu_nfold_by_time();
}

int main(int argc, char **argv) {
  bool cpu = argc > 1 && std::string(argv[1]) == "--cpu";
  bool gpu = argc > 1 && std::string(argv[1]) == "--gpu";
  try {
    if (cpu) {
      cpu_tests();
      cpu_canonical_tests();
      cpu_occupant_tests();
      cpu_kmc_atom_tracking();
      cpu_selected_event_tests();
      cpu_nfold_argument_tests();
    }
    if (gpu) gpu_tests();
  } catch (std::exception const &e) {
    std::cout << "EXCEPTION: " << e.what() << std::endl;
    return 2;
  }
  if (g_fail) {
    std::cout << g_fail << " FAILURES" << std::endl;
    return 1;
  }
  std::cout << "ALL OK" << std::endl;
  return 0;
}
