// JSON input and output of SamplingParams, as libcasm-monte's
// src/casm/monte/sampling/io/json/SamplingParams_json_io.cc:70-229: the keys
// sample_by, spacing, begin, period, base, shift, stochastic_sample_period,
// quantities, json_quantities and sample_trajectory, with the same defaults
// and checks.  Errors are collected and thrown together (the reference
// collects them in an InputParser).
#ifndef CASM_MI355X_monte_sampling_SamplingParams_json_io
#define CASM_MI355X_monte_sampling_SamplingParams_json_io

#include <set>

#include "casm/casm_io/json/jsonParser.hh"
#include "casm/monte/sampling/Sampler.hh"

namespace CASM {
namespace monte {

inline SamplingParams sampling_params_from_json(
    jsonParser const &json, std::set<std::string> const &sampling_function_names,
    std::set<std::string> const &json_sampling_function_names, bool time_sampling_allowed) {
  std::vector<std::string> errors;
  SamplingParams sp;
  if (!json.contains("sample_by")) {
    errors.push_back("sample_by: Required property 'sample_by' not found.");
  } else {
    std::string mode = json["sample_by"].get<std::string>();
    if (mode == "pass")
      sp.sample_mode = SAMPLE_MODE::BY_PASS;
    else if (mode == "step")
      sp.sample_mode = SAMPLE_MODE::BY_STEP;
    else if (time_sampling_allowed && mode == "time")
      sp.sample_mode = SAMPLE_MODE::BY_TIME;
    else if (time_sampling_allowed)
      errors.push_back(
          "sample_by: Error: \"sample_mode\" must be one of \"pass\", \"step\", or \"time\".");
    else
      errors.push_back("sample_by: Error: \"sample_mode\" must be one of \"pass\" or \"step\".");
  }
  std::string method = json.contains("spacing") ? json["spacing"].get<std::string>() : "linear";
  if (method == "linear")
    sp.sample_method = SAMPLE_METHOD::LINEAR;
  else if (method == "log")
    sp.sample_method = SAMPLE_METHOD::LOG;
  else
    errors.push_back("spacing: Error: \"spacing\" must be one of \"linear\", \"log\".");
  sp.begin = json.contains("begin") ? json["begin"].get<double>() : 0.0;
  if (!json.contains("period")) {
    errors.push_back("period: Required property 'period' not found.");
  } else {
    sp.period = json["period"].get<double>();
    if (sp.sample_method == SAMPLE_METHOD::LOG && sp.period <= 1.0)
      errors.push_back("period: Error: For \"spacing\"==\"log\", \"period\" must > 1.0.");
    if (sp.sample_method == SAMPLE_METHOD::LINEAR && sp.period <= 0.0)
      errors.push_back("period: Error: For \"spacing\"==\"log\", \"period\" must > 0.0.");
  }
  sp.base = json.contains("base") ? json["base"].get<double>() : std::pow(10.0, 1.0 / 10.0);
  sp.shift = json.contains("shift") ? json["shift"].get<double>() : 10.0;
  sp.stochastic_sample_period =
      json.contains("stochastic_sample_period") && json["stochastic_sample_period"].get<bool>();
  if (json.contains("quantities"))
    sp.sampler_names = json["quantities"].get<std::vector<std::string>>();
  for (auto const &name : sp.sampler_names)
    if (!sampling_function_names.count(name))
      errors.push_back("quantities: Error: \"" + name + "\" is not a sampling option.");
  if (json.contains("json_quantities"))
    sp.json_sampler_names = json["json_quantities"].get<std::vector<std::string>>();
  for (auto const &name : sp.json_sampler_names)
    if (!json_sampling_function_names.count(name))
      errors.push_back("json_quantities: Error: \"" + name + "\" is not a JSON sampling option.");
  sp.do_sample_trajectory =
      json.contains("sample_trajectory") && json["sample_trajectory"].get<bool>();
  sp.do_sample_time = time_sampling_allowed;
  if (!errors.empty()) {
    std::string msg = "Error reading SamplingParams from JSON:";
    for (auto const &e : errors) msg += "\n  " + e;
    throw std::runtime_error(msg);
  }
  return sp;
}

inline jsonParser &to_json(SamplingParams const &sp, jsonParser &json) {
  json.put_obj();
  if (sp.sample_mode == SAMPLE_MODE::BY_PASS)
    json["sample_by"] = "pass";
  else if (sp.sample_mode == SAMPLE_MODE::BY_STEP)
    json["sample_by"] = "step";
  else
    json["sample_by"] = "time";
  if (sp.sample_method == SAMPLE_METHOD::LINEAR) {
    json["spacing"] = "linear";
    auto integral_or_real = [&](double v, const char *key) {
      if (std::abs(v - std::round(v)) < 1e-5)  // CASM::TOL of CASMcode_global, 1e-5
        json[key] = static_cast<CountType>(std::round(v));
      else
        json[key] = v;
    };
    integral_or_real(sp.begin, "begin");
    integral_or_real(sp.period, "period");
  } else if (sp.sample_method == SAMPLE_METHOD::LOG) {
    json["spacing"] = "log";
    json["begin"] = sp.begin;
    json["base"] = sp.base;
    json["shift"] = sp.shift;
  } else {
    throw std::runtime_error("Error converting SamplingParams to json: invalid sample_method");
  }
  if (sp.stochastic_sample_period) json["stochastic_sample_period"] = true;
  if (!sp.sampler_names.empty()) json["quantities"] = sp.sampler_names;
  if (!sp.json_sampler_names.empty()) json["json_quantities"] = sp.json_sampler_names;
  if (sp.do_sample_trajectory) json["sample_trajectory"] = true;
  return json;
}

}  // namespace monte
}  // namespace CASM

#endif
