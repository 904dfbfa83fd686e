// SamplingFixtureParams from JSON, as libcasm-monte's
// run_management/io/json/SamplingFixtureParams_json_io.hh:37-147: the keys
// "sampling" (SamplingParams), "completion_check" (CompletionCheckParams),
// "analysis": {"functions": [...]}, "results_io" ({"method": "json",
// "kwargs": {"output_dir", "write_observations", "write_trajectory"}},
// ResultsIO_json_io.hh:20-62 and polymorphic_method_json_io.hh) and "log"
// ({"file": "status.json", "frequency_in_s": 600.0}); and to_json of the
// parameters (:149-172).  Errors are exceptions with the reference's messages.
#ifndef CASM_MI355X_monte_run_management_SamplingFixtureParams_json_io
#define CASM_MI355X_monte_run_management_SamplingFixtureParams_json_io

#include "casm/monte/run_management/io/json/jsonResultsIO.hh"
#include "casm/monte/sampling/io/json/SamplingParams_json_io.hh"

namespace CASM {
namespace monte {

template <typename ConfigType>
SamplingFixtureParams<ConfigType, BasicStatistics> sampling_fixture_params_from_json(
    jsonParser const &json, std::string label, StateSamplingFunctionMap const &sampling_functions,
    jsonStateSamplingFunctionMap const &json_sampling_functions,
    ResultsAnalysisFunctionMap<ConfigType, BasicStatistics> const &analysis_functions,
    bool time_sampling_allowed) {
  typedef SamplingFixtureParams<ConfigType, BasicStatistics> params_type;
  std::set<std::string> names, json_names;
  for (auto const &p : sampling_functions) names.insert(p.first);
  for (auto const &p : json_sampling_functions) json_names.insert(p.first);
  if (!json.contains("sampling"))
    throw std::runtime_error("sampling: Required property 'sampling' not found.");
  SamplingParams sp =
      sampling_params_from_json(json["sampling"], names, json_names, time_sampling_allowed);
  if (!json.contains("completion_check"))
    throw std::runtime_error("completion_check: Required property 'completion_check' not found.");
  auto ccp = completion_check_params_from_json(json["completion_check"], sampling_functions);

  std::vector<std::string> analysis_names;
  if (json.contains("analysis") && json["analysis"].contains("functions"))
    analysis_names = json["analysis"]["functions"].get<std::vector<std::string>>();
  for (auto const &name : analysis_names)
    if (!analysis_functions.count(name))
      throw std::runtime_error("analysis/functions: Error: function '" + name +
                               "' not recognized");

  typename params_type::results_io_type results_io = nullptr;
  if (json.contains("results_io") && !json["results_io"].is_null()) {
    auto const &rj = json["results_io"];
    if (!rj.contains("method"))
      throw std::runtime_error("results_io/method: Missing required parameter \"method\".");
    if (!rj["method"].is_string() || rj["method"].get<std::string>() != "json")
      throw std::runtime_error(
          "results_io/method: Parameter \"method\" must be one of: \"json\" (only 1 option "
          "currently)");
    std::string output_dir = "output";
    bool write_observations = false, write_trajectory = false;
    if (rj.contains("kwargs")) {
      auto const &kw = rj["kwargs"];
      if (kw.contains("output_dir")) output_dir = kw["output_dir"].get<std::string>();
      if (kw.contains("write_observations"))
        write_observations = kw["write_observations"].get<bool>();
      if (kw.contains("write_trajectory")) write_trajectory = kw["write_trajectory"].get<bool>();
    }
    jsonResultsIO<Results<ConfigType, BasicStatistics>> io(output_dir, write_trajectory,
                                                           write_observations);
    results_io = io.writer();
  }

  MethodLog method_log;
  if (json.contains("log")) {
    std::string file = "status.json";
    double frequency = 600.0;
    auto const &lj = json["log"];
    if (lj.contains("file")) file = lj["file"].get<std::string>();
    if (lj.contains("frequency_in_s")) frequency = lj["frequency_in_s"].get<double>();
    method_log.log_frequency = frequency;
    method_log.logfile_path = std::filesystem::path(file);
    method_log.reset();
  }
  return params_type(label, sampling_functions, json_sampling_functions, analysis_functions, sp,
                     ccp, analysis_names, results_io, method_log);
}

template <typename ConfigType>
jsonParser &to_json(SamplingFixtureParams<ConfigType, BasicStatistics> const &params,
                    jsonParser &json) {
  json.put_obj();
  to_json(params.sampling_params, json["sampling"]);
  to_json(params.completion_check_params, json["completion_check"]);
  json["analysis"] = jsonParser::object();
  json["analysis"]["functions"] = jsonParser::array();
  for (auto const &name : params.analysis_names) json["analysis"]["functions"].push_back(name);
  // the writer hook keeps no settings to report; the only method is "json"
  if (params.results_io) {
    json["results_io"] = jsonParser::object();
    json["results_io"]["method"] = "json";
  } else {
    json["results_io"].put_null();
  }
  if (!params.method_log.logfile_path.empty()) {
    json["log"] = jsonParser::object();
    json["log"]["file"] = params.method_log.logfile_path.string();
    if (params.method_log.log_frequency.has_value())
      json["log"]["frequency_in_s"] = *params.method_log.log_frequency;
  }
  return json;
}

}  // namespace monte
}  // namespace CASM

#endif
