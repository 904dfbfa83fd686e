// JSON input and output of the completion check, as libcasm-monte's
// checks/io/json/{CompletionCheck,ConvergenceCheck,EquilibrationCheck,
// CutoffCheck}_json_io.hh with src/casm/monte/checks/io/json/*.cc,
// sampling/io/json/Sampler_json_io.cc:39-52 (RequestedPrecision) and
// BasicStatistics.cc:210-214: same keys, nesting and special values
// ("did_not_equilibrate").  Parse errors are exceptions carrying the
// reference's messages (the reference collects them in an InputParser).
#ifndef CASM_MI355X_monte_checks_CompletionCheck_json_io
#define CASM_MI355X_monte_checks_CompletionCheck_json_io

#include <set>

#include "casm/casm_io/json/jsonParser.hh"
#include "casm/monte/checks/CompletionCheck.hh"

namespace CASM {
namespace monte {

inline jsonParser &to_json(BasicStatistics const &stats, jsonParser &json) {
  json.put_obj();
  json["mean"] = stats.mean;
  json["calculated_precision"] = stats.calculated_precision;
  return json;
}

/// adds abs_precision / rel_precision to an object
inline jsonParser &to_json(RequestedPrecision const &value, jsonParser &json) {
  if (json.is_null()) json.put_obj();
  if (!json.is_obj())
    throw std::runtime_error(
        "Error writing RequestedPrecision to json: must write to an existing JSON object");
  if (value.abs_convergence_is_required) json["abs_precision"] = value.abs_precision;
  if (value.rel_convergence_is_required) json["rel_precision"] = value.rel_precision;
  return json;
}

inline jsonParser &to_json(IndividualEquilibrationCheckResult const &value, jsonParser &json) {
  json.put_obj();
  json["is_equilibrated"] = value.is_equilibrated;
  if (value.is_equilibrated)
    json["N_samples_for_equilibration"] = value.N_samples_for_equilibration;
  else
    json["N_samples_for_equilibration"] = "did_not_equilibrate";
  return json;
}

inline jsonParser &to_json(EquilibrationCheckResults const &value, jsonParser &json) {
  json.put_obj();
  json["all_equilibrated"] = value.all_equilibrated;
  if (value.all_equilibrated)
    json["N_samples_for_all_to_equilibrate"] = value.N_samples_for_all_to_equilibrate;
  else
    json["N_samples_for_equilibration"] = "did_not_equilibrate";
  json["individual_results"].put_array();
  for (auto const &p : value.individual_results) {
    jsonParser t;
    to_json(p.second, t);
    t["sampler_name"] = p.first.sampler_name;
    t["component_name"] = p.first.component_name;
    t["component_index"] = p.first.component_index;
    json["individual_results"].push_back(t);
  }
  return json;
}

template <typename StatisticsType>
jsonParser &to_json(IndividualConvergenceCheckResult<StatisticsType> const &value,
                    jsonParser &json) {
  json.put_obj();
  json["is_converged"] = value.is_converged;
  to_json(value.requested_precision, json["requested_precision"]);
  to_json(value.stats, json["stats"]);
  return json;
}

template <typename StatisticsType>
jsonParser &to_json(ConvergenceCheckResults<StatisticsType> const &value, jsonParser &json) {
  json.put_obj();
  json["all_converged"] = value.all_converged;
  json["N_samples_for_statistics"] = value.N_samples_for_statistics;
  json["individual_results"].put_array();
  for (auto const &p : value.individual_results) {
    jsonParser t;
    to_json(p.second, t);
    t["sampler_name"] = p.first.sampler_name;
    t["component_name"] = p.first.component_name;
    t["component_index"] = p.first.component_index;
    json["individual_results"].push_back(t);
  }
  return json;
}

inline jsonParser &to_json(CutoffCheckParams const &params, jsonParser &json) {
  json.put_obj();
  auto minmax = [&](const char *key, auto const &mn, auto const &mx) {
    if (!mn.has_value() && !mx.has_value()) return;
    json[key].put_obj();
    if (mn.has_value()) json[key]["min"] = *mn;
    if (mx.has_value()) json[key]["max"] = *mx;
  };
  minmax("count", params.min_count, params.max_count);
  minmax("time", params.min_time, params.max_time);
  minmax("sample", params.min_sample, params.max_sample);
  minmax("clocktime", params.min_clocktime, params.max_clocktime);
  return json;
}

/// CompletionCheck_json_io.hh:371-417
inline jsonParser &to_json(CompletionCheckParams<BasicStatistics> const &params,
                           jsonParser &json) {
  if (json.is_null()) json.put_obj();
  to_json(params.cutoff_params, json["cutoff"]);
  json["convergence"] = jsonParser::array();
  for (auto const &p : params.requested_precision) {
    jsonParser t = jsonParser::object();
    t["quantity"] = p.first.sampler_name;
    t["component_index"] = std::vector<long>{(long)p.first.component_index};
    t["component_name"] = std::vector<std::string>{p.first.component_name};
    to_json(p.second, t);
    json["convergence"].push_back(t);
  }
  if (!params.log_spacing) {
    json["spacing"] = "linear";
    json["begin"] = params.check_begin;
    json["period"] = params.check_period;
  } else {
    json["spacing"] = "log";
    json["begin"] = params.check_begin;
    json["base"] = params.check_base;
    json["shift"] = params.check_shift;
    json["period_max"] = params.check_period_max;
  }
  return json;
}

/// CompletionCheck_json_io.hh:420-440
template <typename StatisticsType>
jsonParser &to_json(CompletionCheckResults<StatisticsType> const &value, jsonParser &json) {
  json.put_obj();
  json["has_all_minimums_met"] = value.has_all_minimums_met;
  json["has_any_maximum_met"] = value.has_any_maximum_met;
  json["count"] = value.count;
  json["time"] = value.time;
  json["clocktime"] = value.clocktime;
  json["n_samples"] = value.n_samples;
  json["is_complete"] = value.is_complete;
  if (value.n_samples_at_convergence_check.has_value()) {
    json["n_samples_at_convergence_check"] = value.n_samples_at_convergence_check;
    to_json(value.equilibration_check_results, json["equilibration_check_results"]);
    to_json(value.convergence_check_results, json["convergence_check_results"]);
  }
  return json;
}

// --- CompletionCheckParams from JSON (checks/io/json/CompletionCheck_json_io.hh:295-369) --
// Errors are reported by exceptions carrying the reference's messages.

inline CutoffCheckParams cutoff_params_from_json(jsonParser const &json) {
  CutoffCheckParams p;
  auto minmax = [&](std::string const &key, auto &mn, auto &mx) {
    if (!json.contains(key)) return;
    auto const &j = json[key];
    using MinT = typename std::decay<decltype(mn)>::type::value_type;
    if (j.contains("min")) mn = static_cast<MinT>(j["min"].get<double>());
    if (j.contains("max")) mx = static_cast<MinT>(j["max"].get<double>());
  };
  minmax("count", p.min_count, p.max_count);
  minmax("sample", p.min_sample, p.max_sample);
  minmax("time", p.min_time, p.max_time);
  minmax("clocktime", p.min_clocktime, p.max_clocktime);
  return p;
}

inline CompletionCheckParams<BasicStatistics> completion_check_params_from_json(
    jsonParser const &json, StateSamplingFunctionMap const &sampling_functions) {
  CompletionCheckParams<BasicStatistics> p;
  double confidence = json.contains("confidence") ? json["confidence"].get<double>() : 0.95;
  long method = json.contains("weighted_observations_method")
                    ? json["weighted_observations_method"].get<long>()
                    : 1;
  long n_resamples = json.contains("n_resamples") ? json["n_resamples"].get<long>() : 10000;
  p.equilibration_check_f = default_equilibration_check;
  p.calc_statistics_f = BasicStatisticsCalculator(confidence, method, n_resamples);
  if (json.contains("cutoff")) p.cutoff_params = cutoff_params_from_json(json["cutoff"]);
  if (json.contains("convergence")) {
    for (auto const &c : json["convergence"].elements()) {
      std::string q = c["quantity"].get<std::string>();
      auto it = sampling_functions.find(q);
      if (it == sampling_functions.end())
        throw std::runtime_error("Error: \"" + q + "\" is not a sampling option.");
      RequestedPrecision rp;
      if (c.contains("precision")) {
        rp.abs_convergence_is_required = true;
        rp.abs_precision = c["precision"].get<double>();
      }
      if (c.contains("abs_precision")) {
        rp.abs_convergence_is_required = true;
        rp.abs_precision = c["abs_precision"].get<double>();
      }
      if (c.contains("rel_precision")) {
        rp.rel_convergence_is_required = true;
        rp.rel_precision = c["rel_precision"].get<double>();
      }
      auto const &names = it->second.component_names;
      std::set<Index> selected;
      if (c.contains("component_index"))
        for (long i : c["component_index"].get<std::vector<long>>()) selected.insert(i);
      if (c.contains("component_name"))
        for (auto const &n : c["component_name"].get<std::vector<std::string>>()) {
          auto f = std::find(names.begin(), names.end(), n);
          if (f == names.end())
            throw std::runtime_error("Error: \"" + n + "\" is not a component of \"" + q + "\"");
          selected.insert(f - names.begin());
        }
      if (selected.empty())
        for (Index i = 0; i < (Index)names.size(); ++i) selected.insert(i);
      for (Index i : selected) {
        if (i < 0 || i >= (Index)names.size())
          throw std::runtime_error("Error: component_index out of range for \"" + q + "\"");
        p.requested_precision[SamplerComponent(q, i, names[i])] = rp;
      }
    }
  }
  std::string spacing = json.contains("spacing") ? json["spacing"].get<std::string>() : "linear";
  if (spacing == "linear") {
    p.log_spacing = false;
    p.check_begin = 100;
    p.check_period = 100;
  } else if (spacing == "log") {
    p.log_spacing = true;
    p.check_begin = 0;
    p.check_base = 10.0;
    p.check_shift = 2.0;
    p.check_period_max = 10000;
  } else {
    throw std::runtime_error("Error: \"spacing\" must be one of \"linear\", \"log\".");
  }
  if (json.contains("begin")) p.check_begin = json["begin"].get<long>();
  if (json.contains("period")) p.check_period = json["period"].get<long>();
  if (p.check_period <= 1) throw std::runtime_error("Error: \"period\" must > 0.");
  if (json.contains("base")) p.check_base = json["base"].get<double>();
  if (json.contains("shift")) p.check_shift = json["shift"].get<double>();
  if (json.contains("period_max")) p.check_period_max = json["period_max"].get<long>();
  if (p.check_base <= 1.0) throw std::runtime_error("Error: \"base\" must > 1.0");
  return p;
}

}  // namespace monte
}  // namespace CASM

#endif
