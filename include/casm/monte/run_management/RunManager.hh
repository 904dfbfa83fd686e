// Drop-in counterparts of libcasm-monte's run management:
//  State (run_management/State.hh:15-42), Results (Results.hh:18-159),
//  ResultsAnalysisFunction / make_analysis (ResultsAnalysisFunction.hh:20-136),
//  SamplingFixtureParams, MonteCounter, SamplingFixture (SamplingFixture.hh:27-644),
//  RunManager (RunManager.hh:19-236).
// Same members, call order and decisions; results I/O is a callback hook
// instead of the reference's JSON writer.
#ifndef CASM_MI355X_monte_RunManager
#define CASM_MI355X_monte_RunManager

#include "casm/monte/checks/CompletionCheck.hh"
#include "casm/monte/checks/io/json/CompletionCheck_json_io.hh"
#include "casm/monte/sampling/SelectedEventFunctions.hh"

namespace CASM {
namespace monte {

template <typename _ConfigType>
struct State {
  typedef _ConfigType ConfigType;
  explicit State(ConfigType const &_configuration, ValueMap _conditions = ValueMap(),
                 ValueMap _properties = ValueMap())
      : configuration(_configuration), conditions(_conditions), properties(_properties) {}
  ConfigType configuration;
  ValueMap conditions;
  ValueMap properties;
};

template <typename ConfigType, typename StatisticsType>
struct Results;

template <typename ConfigType, typename StatisticsType>
struct ResultsAnalysisFunction {
  ResultsAnalysisFunction(
      std::string _name, std::string _description, std::vector<Index> _shape,
      std::function<VectorXd(Results<ConfigType, StatisticsType> const &)> _function)
      : name(_name),
        description(_description),
        shape(_shape),
        component_names(default_component_names(_shape)),
        function(_function) {}
  std::string name, description;
  std::vector<Index> shape;
  std::vector<std::string> component_names;
  std::function<VectorXd(Results<ConfigType, StatisticsType> const &)> function;
  VectorXd operator()(Results<ConfigType, StatisticsType> const &r) const { return function(r); }
};

template <typename ConfigType, typename StatisticsType>
using ResultsAnalysisFunctionMap =
    std::map<std::string, ResultsAnalysisFunction<ConfigType, StatisticsType>>;

/// Resident memory of the process in MiB (0 when /proc is unavailable); the
/// reference records it at the start and end of every run (Results.hh).
inline double memory_used_MiB() {
  std::ifstream f("/proc/self/status");
  std::string line;
  while (std::getline(f, line))
    if (line.rfind("VmRSS:", 0) == 0) return std::stod(line.substr(6)) / 1024.0;
  return 0.0;
}

template <typename ConfigType, typename StatisticsType>
struct Results {
  typedef ConfigType config_type;
  typedef StatisticsType stats_type;
  Results(std::vector<std::string> _sampler_names,
          StateSamplingFunctionMap const &_sampling_functions,
          std::vector<std::string> _json_sampler_names,
          jsonStateSamplingFunctionMap const &_json_sampling_functions,
          ResultsAnalysisFunctionMap<ConfigType, StatisticsType> const &_analysis_functions)
      : sampler_names(_sampler_names),
        sampling_functions(_sampling_functions),
        json_sampler_names(_json_sampler_names),
        json_sampling_functions(_json_sampling_functions),
        analysis_functions(_analysis_functions),
        sample_weight({}) {}
  std::vector<std::string> sampler_names;
  StateSamplingFunctionMap sampling_functions;
  std::vector<std::string> json_sampler_names;
  jsonStateSamplingFunctionMap json_sampling_functions;
  ResultsAnalysisFunctionMap<ConfigType, StatisticsType> analysis_functions;
  std::optional<TimeType> elapsed_clocktime;
  double initial_memory_used_MiB = 0.0;
  double final_memory_used_MiB = 0.0;
  SamplerMap samplers;
  std::map<std::string, std::shared_ptr<jsonSampler>> json_samplers;
  std::map<std::string, VectorXd> analysis;
  std::vector<CountType> sample_count;
  std::vector<TimeType> sample_time;
  Sampler sample_weight;
  std::vector<TimeType> sample_clocktime;
  std::vector<ConfigType> sample_trajectory;
  CompletionCheckResults<StatisticsType> completion_check_results;
  BigCountType n_accept = 0;
  BigCountType n_reject = 0;
  /// selected-event data a kinetic run collected (Results.hh:73-74)
  std::optional<SelectedEventData> selected_event_data;

  void reset() {
    elapsed_clocktime.reset();
    selected_event_data.reset();
    initial_memory_used_MiB = memory_used_MiB();
    final_memory_used_MiB = 0.0;
    samplers.clear();
    json_samplers.clear();
    analysis.clear();
    sample_count.clear();
    sample_time.clear();
    sample_weight.clear();
    sample_clocktime.clear();
    sample_trajectory.clear();
    completion_check_results.full_reset();
    n_accept = 0;
    n_reject = 0;
    for (auto const &name : sampler_names) {
      auto it = sampling_functions.find(name);
      if (it == sampling_functions.end())
        throw std::runtime_error("Results::reset error. Failed to find sampling function '" +
                                 name + "'.");
      samplers.emplace(it->second.name, std::make_shared<Sampler>(it->second.shape,
                                                                  it->second.component_names));
    }
    for (auto const &name : json_sampler_names) {
      auto it = json_sampling_functions.find(name);
      if (it == json_sampling_functions.end())
        throw std::runtime_error(
            "Results::reset error. Failed to find json sampling function '" + name + "'.");
      json_samplers.emplace(it->second.name, std::make_shared<jsonSampler>());
    }
  }
};

template <typename ConfigType, typename StatisticsType>
std::map<std::string, VectorXd> make_analysis(
    Results<ConfigType, StatisticsType> const &results,
    ResultsAnalysisFunctionMap<ConfigType, StatisticsType> const &functions,
    std::vector<std::string> names) {
  std::map<std::string, VectorXd> analysis;
  for (auto const &name : names) {
    auto it = functions.find(name);
    if (it == functions.end()) continue;
    auto const &f = it->second;
    try {
      analysis.emplace(f.name, f(results));
    } catch (std::exception const &e) {
      std::cerr << "Results analysis '" << it->first << "' failed: " << e.what() << std::endl;
      analysis.emplace(f.name, VectorXd::Constant((Index)f.component_names.size(),
                                                  std::numeric_limits<double>::signaling_NaN()));
    }
  }
  return analysis;
}

template <typename ConfigType, typename StatisticsType>
struct SamplingFixtureParams {
  typedef Results<ConfigType, StatisticsType> results_type;
  /// Called by finalize with (results, conditions, run_index): the results
  /// writer hook (the reference's ResultsIO::write).
  typedef std::function<void(results_type const &, ValueMap const &, Index)> results_io_type;

  SamplingFixtureParams(
      std::string _label, StateSamplingFunctionMap _sampling_functions,
      jsonStateSamplingFunctionMap _json_sampling_functions,
      ResultsAnalysisFunctionMap<ConfigType, StatisticsType> _analysis_functions,
      SamplingParams _sampling_params,
      CompletionCheckParams<StatisticsType> _completion_check_params,
      std::vector<std::string> _analysis_names, results_io_type _results_io = nullptr,
      MethodLog _method_log = MethodLog())
      : label(_label),
        sampling_functions(_sampling_functions),
        json_sampling_functions(_json_sampling_functions),
        analysis_functions(_analysis_functions),
        sampling_params(_sampling_params),
        completion_check_params(_completion_check_params),
        analysis_names(_analysis_names),
        results_io(_results_io),
        method_log(_method_log) {
    for (auto const &name : sampling_params.sampler_names)
      if (!sampling_functions.count(name))
        throw std::runtime_error(
            "SamplingFixtureParams constructor error: No sampling function for '" + name + "'");
    for (auto const &name : sampling_params.json_sampler_names)
      if (!json_sampling_functions.count(name))
        throw std::runtime_error(
            "SamplingFixtureParams constructor error: No sampling function for '" + name + "'");
  }
  std::string label;
  StateSamplingFunctionMap sampling_functions;
  jsonStateSamplingFunctionMap json_sampling_functions;
  ResultsAnalysisFunctionMap<ConfigType, StatisticsType> analysis_functions;
  SamplingParams sampling_params;
  CompletionCheckParams<StatisticsType> completion_check_params;
  std::vector<std::string> analysis_names;
  results_io_type results_io;
  MethodLog method_log;
};

struct MonteCounter {
  MonteCounter() { reset(SAMPLE_MODE::BY_PASS, 1); }
  SAMPLE_MODE sample_mode;
  CountType steps_per_pass;
  CountType step, pass, count;
  TimeType time;
  BigCountType n_accept, n_reject;
  void reset(SAMPLE_MODE m, CountType spp) {
    sample_mode = m;
    steps_per_pass = spp;
    step = pass = count = 0;
    time = 0.0;
    n_accept = n_reject = 0;
  }
  void increment_n_accept() { ++n_accept; }
  void increment_n_reject() { ++n_reject; }
  void increment_step() {
    ++step;
    if (sample_mode == SAMPLE_MODE::BY_STEP) ++count;
    if (step == steps_per_pass) {
      ++pass;
      if (sample_mode != SAMPLE_MODE::BY_STEP) ++count;
      step = 0;
    }
  }
  /// Advance by whole passes (the GPU loop's counterpart of repeated
  /// increment_step calls; only valid at a pass boundary).
  void increment_passes(CountType n_passes, BigCountType accepted, BigCountType rejected) {
    pass += n_passes;
    count += (sample_mode == SAMPLE_MODE::BY_STEP) ? n_passes * steps_per_pass : n_passes;
    n_accept += accepted;
    n_reject += rejected;
  }
  /// Advance by k steps from any position (k calls of increment_step).
  void increment_steps(CountType k, BigCountType accepted, BigCountType rejected) {
    const CountType s = step + k;
    const CountType passes = s / steps_per_pass;
    pass += passes;
    step = s - passes * steps_per_pass;
    count += (sample_mode == SAMPLE_MODE::BY_STEP) ? k : passes;
    n_accept += accepted;
    n_reject += rejected;
  }
  void set_time(double t) { time = t; }
};

template <typename _ConfigType, typename _StatisticsType, typename _EngineType>
class SamplingFixture {
 public:
  typedef _ConfigType config_type;
  typedef _StatisticsType stats_type;
  typedef _EngineType engine_type;
  typedef State<config_type> state_type;

  SamplingFixture(SamplingFixtureParams<config_type, stats_type> const &params,
                  std::shared_ptr<engine_type> engine)
      : m_params(params),
        m_engine(engine),
        m_random_number_generator(engine),
        m_completion_check(m_params.completion_check_params),
        m_results(m_params.sampling_params.sampler_names, m_params.sampling_functions,
                  m_params.sampling_params.json_sampler_names, m_params.json_sampling_functions,
                  m_params.analysis_functions) {}

  std::string label() const { return m_params.label; }
  SamplingFixtureParams<config_type, stats_type> const &params() const { return m_params; }
  MonteCounter const &counter() const { return m_counter; }
  MonteCounter &counter() { return m_counter; }
  Results<config_type, stats_type> const &results() const { return m_results; }
  CompletionCheck<stats_type> const &completion_check() const { return m_completion_check; }

  void initialize(Index steps_per_pass) {
    m_n_samples = 0;
    m_count = 0;
    m_is_complete = false;
    m_counter.reset(m_params.sampling_params.sample_mode, steps_per_pass);
    m_completion_check.reset();
    m_results.reset();
    // SamplingFixture.hh:219-235
    if (m_params.sampling_params.sample_mode == SAMPLE_MODE::BY_TIME) {
      m_next_sample_count = 0;
      m_next_sample_time = sample_at(m_results.sample_time.size());
      if (m_next_sample_time < 0.0)
        throw std::runtime_error("Error: sampling period parameter error, next_sample_time < 0.0");
    } else {
      m_next_sample_time = 0.0;
      m_next_sample_count =
          static_cast<CountType>(std::round(sample_at(m_results.sample_count.size())));
      if (m_next_sample_count < 0)
        throw std::runtime_error("Error: sampling period parameter error, next_sample_count < 0");
    }
    Log &log = m_params.method_log.log;
    log.restart_clock();
    log.begin_lap();
  }

  bool is_complete() {
    if (m_is_complete) return true;
    Log &log = m_params.method_log.log;
    // SamplingFixture.hh:271-279: fixtures that sample time also hand the
    // counter's time to the check (its results then record it)
    if (m_params.sampling_params.do_sample_time)
      m_is_complete = m_completion_check.is_complete(m_results.samplers, m_results.sample_weight,
                                                     m_counter.count, m_counter.time, log);
    else
      m_is_complete = m_completion_check.is_complete(m_results.samplers, m_results.sample_weight,
                                                     m_counter.count, log);
    return m_is_complete;
  }

  void write_status(Index run_index) {
    if (m_params.method_log.logfile_path.empty()) return;
    Log &log = m_params.method_log.log;
    m_params.method_log.reset();
    jsonParser json = jsonParser::object();
    json["run_index"] = run_index;
    json["time"] = log.time_s();
    to_json(m_completion_check.results(), json["completion_check_results"]);
    std::ostringstream os;
    json.print(os);
    log << os.str() << std::endl;
    log.begin_lap();
  }

  void write_status_if_due(Index run_index) {
    std::optional<double> &f = m_params.method_log.log_frequency;
    if (!f.has_value()) return;
    if (m_n_samples != get_n_samples(m_results.samplers) || m_count != m_counter.count) {
      m_n_samples = get_n_samples(m_results.samplers);
      m_count = m_counter.count;
      if (m_params.method_log.log.lap_time() > *f) write_status(run_index);
    }
  }

  void increment_n_accept() { m_counter.increment_n_accept(); }
  void increment_n_reject() { m_counter.increment_n_reject(); }
  void increment_step() { m_counter.increment_step(); }
  CountType next_sample_count() const { return m_next_sample_count; }
  TimeType next_sample_time() const { return m_next_sample_time; }

  /// Weight of the next sample (time increment of the N-fold way's event
  /// selected at the sample; SamplingFixture.hh:330-332)
  void push_back_sample_weight(double weight) { m_results.sample_weight.push_back(weight); }
  /// store the selected-event data a run collected (SamplingFixture.hh:340-343)
  void set_selected_event_data(SelectedEventData &&selected_event_data) {
    m_results.selected_event_data = std::move(selected_event_data);
  }
  void set_time(double t) { m_counter.set_time(t); }

  void sample_data(state_type const &state) {
    m_results.sample_count.push_back(m_counter.count);
    if (m_params.sampling_params.do_sample_time) m_results.sample_time.push_back(m_counter.time);
    m_results.sample_clocktime.push_back(m_params.method_log.log.time_s());
    if (m_params.sampling_params.do_sample_trajectory)
      m_results.sample_trajectory.push_back(state.configuration);
    for (auto const &name : m_params.sampling_params.sampler_names)
      m_results.samplers.at(name)->push_back(m_params.sampling_functions.at(name)());
    for (auto const &name : m_params.sampling_params.json_sampler_names)
      m_results.json_samplers.at(name)->values.push_back(
          m_params.json_sampling_functions.at(name)());
    // SamplingFixture.hh:516-538
    if (m_params.sampling_params.sample_mode == SAMPLE_MODE::BY_TIME) {
      m_next_sample_time = sample_at(m_results.sample_time.size());
      if (m_next_sample_time <= m_counter.time)
        throw std::runtime_error(
            "Error: state sampling period parameter error, next_sample_time <= current time");
    } else {
      m_next_sample_count =
          static_cast<CountType>(std::round(sample_at(m_results.sample_count.size())));
      if (m_next_sample_count <= m_counter.count)
        throw std::runtime_error(
            "Error: state sampling period parameter error, next_sample_count <= current count");
    }
  }

  void sample_data_by_count_if_due(state_type const &state) {
    if (m_params.sampling_params.sample_mode != SAMPLE_MODE::BY_TIME &&
        m_counter.count == m_next_sample_count)
      sample_data(state);
  }

  /// SamplingFixture.hh:570-588: a stochastic period draws the next sample's
  /// count from the fixture's generator, which shares the run's engine
  double sample_at(CountType sample_index) {
    if (m_params.sampling_params.stochastic_sample_period)
      return stochastic_sample_at(sample_index, m_params.sampling_params,
                                  m_random_number_generator, m_results.sample_count,
                                  m_results.sample_time);
    return monte::sample_at(sample_index, m_params.sampling_params);
  }

  void finalize(state_type const &state, Index run_index) {
    Log &log = m_params.method_log.log;
    m_results.elapsed_clocktime = log.time_s();
    m_results.final_memory_used_MiB = memory_used_MiB();
    m_results.completion_check_results = m_completion_check.results();
    m_results.analysis =
        make_analysis(m_results, m_params.analysis_functions, m_params.analysis_names);
    m_results.n_accept = m_counter.n_accept;
    m_results.n_reject = m_counter.n_reject;
    if (m_params.results_io) m_params.results_io(m_results, state.conditions, run_index);
    write_status(run_index);
  }

 private:
  SamplingFixtureParams<config_type, stats_type> m_params;
  std::shared_ptr<engine_type> m_engine;
  RandomNumberGenerator<engine_type> m_random_number_generator;
  Index m_n_samples = 0;
  Index m_count = 0;
  bool m_is_complete = false;
  MonteCounter m_counter;
  CountType m_next_sample_count = 0;
  TimeType m_next_sample_time = 0.0;
  CompletionCheck<stats_type> m_completion_check;
  Results<config_type, stats_type> m_results;
};

template <typename _ConfigType, typename _StatisticsType, typename _EngineType>
struct RunManager {
  typedef _ConfigType config_type;
  typedef _StatisticsType stats_type;
  typedef _EngineType engine_type;
  typedef State<config_type> state_type;
  typedef SamplingFixtureParams<config_type, stats_type> sampling_fixture_params_type;
  typedef SamplingFixture<config_type, stats_type, engine_type> sampling_fixture_type;

  Index run_index = 0;
  std::shared_ptr<engine_type> engine;
  std::vector<std::shared_ptr<sampling_fixture_type>> sampling_fixtures;
  bool global_cutoff;

  RunManager(std::shared_ptr<engine_type> _engine,
             std::vector<sampling_fixture_params_type> const &params, bool _global_cutoff = true)
      : engine(_engine), global_cutoff(_global_cutoff) {
    if (!engine) throw std::runtime_error("Error constructing RunManager: engine==nullptr");
    for (auto const &p : params)
      sampling_fixtures.emplace_back(std::make_shared<sampling_fixture_type>(p, engine));
  }

  void initialize(Index steps_per_pass) {
    for (auto &f : sampling_fixtures) f->initialize(steps_per_pass);
  }

  bool is_complete() {
    bool all_complete = true, any_complete = false;
    for (auto &f : sampling_fixtures) {
      if (f->is_complete())
        any_complete = true;
      else
        all_complete = false;
    }
    if (global_cutoff && any_complete) return true;
    return all_complete;
  }

  void write_status_if_due() {
    for (auto &f : sampling_fixtures) f->write_status_if_due(run_index);
  }
  void increment_n_accept() {
    for (auto &f : sampling_fixtures) f->increment_n_accept();
  }
  void increment_n_reject() {
    for (auto &f : sampling_fixtures) f->increment_n_reject();
  }
  void increment_step() {
    for (auto &f : sampling_fixtures) f->increment_step();
  }
  struct NullAction {
    void operator()(sampling_fixture_type const &, state_type const &) {}
  };

  void set_time(double event_time) {
    for (auto &f : sampling_fixtures) f->set_time(event_time);
  }

  /// Sample every fixture due by count (RunManager.hh:165-186); the actions
  /// run just before and after each fixture's sample.
  template <typename PreSampleActionType = NullAction, typename PostSampleActionType = NullAction>
  void sample_data_by_count_if_due(state_type const &state,
                                   PreSampleActionType pre_sample_f = PreSampleActionType(),
                                   PostSampleActionType post_sample_f = PostSampleActionType()) {
    for (auto &f : sampling_fixtures)
      if (f->params().sampling_params.sample_mode != SAMPLE_MODE::BY_TIME &&
          f->counter().count == f->next_sample_count()) {
        pre_sample_f(*f, state);
        f->sample_data(state);
        post_sample_f(*f, state);
      }
  }
  /// Sample every fixture due by time, earliest first, at its sample time
  /// (RunManager.hh:190-209): the fixture's time is set to its sample time
  /// before it samples.
  template <typename PreSampleActionType = NullAction, typename PostSampleActionType = NullAction>
  void sample_data_by_time_if_due(TimeType event_time, state_type const &state,
                                  PreSampleActionType pre_sample_f = PreSampleActionType(),
                                  PostSampleActionType post_sample_f = PostSampleActionType()) {
    while (next_sampling_fixture != nullptr && event_time >= next_sample_time) {
      auto &fixture = *next_sampling_fixture;
      pre_sample_f(fixture, state);
      fixture.set_time(next_sample_time);
      fixture.sample_data(state);
      post_sample_f(fixture, state);
      update_next_sampling_fixture();
    }
  }

  /// The BY_TIME fixture that samples next and its sample time
  /// (RunManager.hh:211-225)
  void update_next_sampling_fixture() {
    next_sampling_fixture = nullptr;
    for (auto &f : sampling_fixtures)
      if (f->params().sampling_params.sample_mode == SAMPLE_MODE::BY_TIME &&
          (next_sampling_fixture == nullptr || f->next_sample_time() < next_sample_time)) {
        next_sample_time = f->next_sample_time();
        next_sampling_fixture = f.get();
      }
  }
  sampling_fixture_type *next_sampling_fixture = nullptr;
  TimeType next_sample_time = 0.0;

  void finalize(state_type const &final_state) {
    for (auto &f : sampling_fixtures) f->finalize(final_state, run_index);
  }
};

}  // namespace monte
}  // namespace CASM

#endif
