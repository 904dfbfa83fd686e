// JSON input and output of the selected-event statistics, as libcasm-monte's
// src/casm/monte/sampling/io/json/SelectedEventFunctions_json_io.cc: the
// parameters (correlations_data_params, quantities and the per-quantity tol,
// bin_width, initial_begin, spacing and max_size maps), the hop-correlation
// data and every histogram type, with the same keys and nesting so that files
// written by one are read by the other.  Matrices are written as arrays of
// rows.  Errors are collected and thrown together.
#ifndef CASM_MI355X_monte_sampling_SelectedEventFunctions_json_io
#define CASM_MI355X_monte_sampling_SelectedEventFunctions_json_io

#include "casm/casm_io/json/jsonParser.hh"
#include "casm/monte/sampling/SelectedEventFunctions.hh"

namespace CASM {
namespace monte {

namespace selected_event_json {

inline jsonParser rows(MatrixXl const &M, Index n_rows) {
  jsonParser j = jsonParser::array();
  for (Index i = 0; i < n_rows; ++i) {
    std::vector<long> r;
    for (Index k = 0; k < M.cols(); ++k) r.push_back(M(i, k));
    j.push_back(jsonParser(r));
  }
  return j;
}

inline jsonParser rows(MatrixXd const &M, Index n_rows) {
  jsonParser j = jsonParser::array();
  for (Index i = 0; i < n_rows; ++i) {
    std::vector<double> r;
    for (Index k = 0; k < M.cols(); ++k) r.push_back(M(i, k));
    j.push_back(jsonParser(r));
  }
  return j;
}

inline jsonParser value(VectorXl const &v) { return jsonParser(v.vector()); }
inline jsonParser value(VectorXd const &v) { return jsonParser(v.vector()); }

template <typename T>
jsonParser map_to_json(std::map<std::string, T> const &m) {
  jsonParser j = jsonParser::object();
  for (auto const &pr : m) j[pr.first] = jsonParser(pr.second);
  return j;
}

}  // namespace selected_event_json

/// CorrelationsDataParams (SelectedEventFunctions_json_io.cc:12-40)
inline jsonParser &to_json(CorrelationsDataParams const &params, jsonParser &json) {
  json.put_obj();
  json["jumps_per_position_sample"] = params.jumps_per_position_sample;
  json["max_n_position_samples"] = params.max_n_position_samples;
  json["output_incomplete_samples"] = params.output_incomplete_samples;
  json["stop_run_when_complete"] = params.stop_run_when_complete;
  return json;
}

inline CorrelationsDataParams correlations_data_params_from_json(jsonParser const &json) {
  CorrelationsDataParams p;
  if (json.contains("jumps_per_position_sample"))
    p.jumps_per_position_sample = json["jumps_per_position_sample"].get<long>();
  if (json.contains("max_n_position_samples"))
    p.max_n_position_samples = json["max_n_position_samples"].get<long>();
  if (json.contains("output_incomplete_samples"))
    p.output_incomplete_samples = json["output_incomplete_samples"].get<bool>();
  if (json.contains("stop_run_when_complete"))
    p.stop_run_when_complete = json["stop_run_when_complete"].get<bool>();
  return p;
}

/// CorrelationsData (SelectedEventFunctions_json_io.cc:42-74): the complete
/// position samples only, unless output_incomplete_samples
inline jsonParser &to_json(CorrelationsData const &data, jsonParser &json) {
  using namespace selected_event_json;
  json.put_obj();
  json["jumps_per_position_sample"] = data.jumps_per_position_sample;
  json["max_n_position_samples"] = data.max_n_position_samples;
  json["output_incomplete_samples"] = data.output_incomplete_samples;
  std::vector<long> nps(data.n_position_samples.begin(), data.n_position_samples.end());
  json["n_position_samples"] = jsonParser(nps);
  json["n_complete_samples"] = data.n_complete_samples;
  const Index n = data.output_incomplete_samples ? data.step.rows() : data.n_complete_samples;
  json["step"] = rows(data.step, n);
  json["pass"] = rows(data.pass, n);
  json["time"] = rows(data.time, n);
  json["atom_positions_cart"] = jsonParser::array();
  for (Index i = 0; i < n; ++i)
    json["atom_positions_cart"].push_back(rows(data.atom_positions_cart[i], 3));
  return json;
}

/// Discrete histograms (SelectedEventFunctions_json_io.cc:76-140)
template <typename ValueType, typename CompareType>
jsonParser &to_json(DiscreteVectorHistogramT<ValueType, CompareType> const &data,
                    jsonParser &json) {
  using namespace selected_event_json;
  json.put_obj();
  std::vector<long> shape(data.shape().begin(), data.shape().end());
  json["shape"] = jsonParser(shape);
  json["max_size"] = data.max_size();
  json["max_size_exceeded"] = data.max_size_exceeded();
  json["size"] = data.size();
  json["sum"] = data.sum();
  if (!data.shape().empty()) json["component_names"] = jsonParser(data.component_names());
  if (data.value_labels().has_value()) {
    json["labels"] = jsonParser::array();
    auto const &labels = *data.value_labels();
    for (auto const &v : data.values()) {
      if (labels.count(v))
        json["labels"].push_back(jsonParser(labels.at(v)));
      else if (data.shape().empty())
        json["labels"].push_back(jsonParser(v(0)));
      else
        json["labels"].push_back(value(v));
    }
  }
  json["values"] = jsonParser::array();
  for (auto const &v : data.values())
    json["values"].push_back(data.shape().empty() ? jsonParser(v(0)) : value(v));
  json["count"] = jsonParser(data.count());
  json["fraction"] = jsonParser(data.fraction());
  json["out_of_range_count"] = data.out_of_range_count();
  return json;
}

/// Histogram1D (SelectedEventFunctions_json_io.cc:142-158)
inline jsonParser &to_json(Histogram1D const &data, jsonParser &json) {
  json.put_obj();
  json["max_size"] = data.max_size();
  json["max_size_exceeded"] = data.max_size_exceeded();
  json["is_log"] = data.is_log();
  json["begin"] = data.begin();
  json["bin_width"] = data.bin_width();
  json["count"] = jsonParser(data.count());
  json["bin_coords"] = jsonParser(data.bin_coords());
  json["sum"] = data.sum();
  json["density"] = jsonParser(data.density());
  json["out_of_range_count"] = data.out_of_range_count();
  return json;
}

/// PartitionedHistogram1D (SelectedEventFunctions_json_io.cc:160-188): one
/// partition as a plain histogram; several as the combined histogram and
/// each partition with its fraction and partial density
inline jsonParser &to_json(PartitionedHistogram1D const &data, jsonParser &json) {
  json.put_obj();
  if (data.partition_names().size() == 1) {
    to_json(data.histograms()[0], json);
    return json;
  }
  to_json(data.combined_histogram(), json["combined"]);
  json["partition_names"] = jsonParser(data.partition_names());
  json["partitions"] = jsonParser::array();
  Index i = 0;
  for (auto const &hist : data.histograms()) {
    const double frac = hist.sum() / data.combined_histogram().sum();
    jsonParser t;
    to_json(hist, t);
    t["partition_name"] = data.partition_names()[i];
    t["partition_fraction"] = frac;
    std::vector<double> partial;
    for (double x : hist.density()) partial.push_back(x * frac);
    t["partial_density"] = jsonParser(partial);
    json["partitions"].push_back(t);
    ++i;
  }
  return json;
}

/// SelectedEventFunctionParams input (SelectedEventFunctions_json_io.cc:
/// 190-247): "quantities" and, per quantity, optional "tol", "bin_width",
/// "initial_begin", "spacing" ("log" or "linear") and "max_size"
inline SelectedEventFunctionParams selected_event_function_params_from_json(
    jsonParser const &json) {
  std::vector<std::string> errors;
  SelectedEventFunctionParams params;
  if (json.contains("correlations_data_params"))
    params.correlations_data_params =
        correlations_data_params_from_json(json["correlations_data_params"]);
  if (json.contains("quantities"))
    params.function_names = json["quantities"].get<std::vector<std::string>>();
  auto opt = [&](std::string const &key, std::string const &name) -> jsonParser const * {
    if (json.contains(key) && json[key].contains(name)) return &json[key][name];
    return nullptr;
  };
  for (std::string const &name : params.function_names) {
    if (auto j = opt("tol", name)) params.tol[name] = j->get<double>();
    if (auto j = opt("bin_width", name)) params.bin_width[name] = j->get<double>();
    if (auto j = opt("initial_begin", name)) params.initial_begin[name] = j->get<double>();
    if (auto j = opt("spacing", name)) {
      const std::string spacing = j->get<std::string>();
      if (spacing == "log")
        params.is_log[name] = true;
      else if (spacing == "linear")
        params.is_log[name] = false;
      else
        errors.push_back("spacing/" + name + ": Error: \"" + spacing +
                         "\" is not a valid spacing option for \"" + name + "\".");
    }
    if (auto j = opt("max_size", name)) params.max_size[name] = j->get<long>();
  }
  if (!errors.empty()) {
    std::string msg = "Error reading SelectedEventFunctionParams from JSON:";
    for (auto const &e : errors) msg += "\n  " + e;
    throw std::runtime_error(msg);
  }
  return params;
}

/// SelectedEventFunctionParams output (SelectedEventFunctions_json_io.cc:
/// 249-283)
inline jsonParser &to_json(SelectedEventFunctionParams const &params, jsonParser &json) {
  using namespace selected_event_json;
  json.put_obj();
  if (params.correlations_data_params.has_value())
    to_json(*params.correlations_data_params, json["correlations_data_params"]);
  json["quantities"] = jsonParser(params.function_names);
  if (!params.tol.empty()) json["tol"] = map_to_json(params.tol);
  if (!params.bin_width.empty()) json["bin_width"] = map_to_json(params.bin_width);
  if (!params.initial_begin.empty()) json["initial_begin"] = map_to_json(params.initial_begin);
  if (!params.is_log.empty()) {
    json["spacing"] = jsonParser::object();
    for (auto const &pr : params.is_log) json["spacing"][pr.first] = pr.second ? "log" : "linear";
  }
  if (!params.max_size.empty()) json["max_size"] = map_to_json(params.max_size);
  return json;
}

/// SelectedEventData (SelectedEventFunctions_json_io.cc:285-313)
inline jsonParser &to_json(SelectedEventData const &data, jsonParser &json) {
  json.put_obj();
  if (data.correlations_data.has_value())
    to_json(*data.correlations_data, json["correlations_data"]);
  const size_t n = data.discrete_vector_int_histograms.size() +
                   data.discrete_vector_float_histograms.size() +
                   data.continuous_1d_histograms.size();
  if (n > 0) {
    json["histograms"] = jsonParser::object();
    for (auto const &pr : data.discrete_vector_int_histograms)
      to_json(pr.second, json["histograms"][pr.first]);
    for (auto const &pr : data.discrete_vector_float_histograms)
      to_json(pr.second, json["histograms"][pr.first]);
    for (auto const &pr : data.continuous_1d_histograms)
      to_json(pr.second, json["histograms"][pr.first]);
  }
  return json;
}

}  // namespace monte
}  // namespace CASM

#endif
