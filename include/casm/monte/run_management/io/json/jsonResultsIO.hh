// jsonResultsIO.hh -- results files of a sampling fixture in the reference's
// formats (run_management/io/json/jsonResultsIO_impl.hh:360-526):
//   <output_dir>/summary.json            one entry appended per completed run
//   <output_dir>/run.<i>/observations.json  the sampled series of run i
//   <output_dir>/run.<i>/trajectory.json    the sampled configurations of run i
// and the per-quantity statistics of run_management/Results.hh:165-352
// (QuantityStats).  Plug it into a SamplingFixtureParams with
// `params.results_io = results_io.writer();`.
#ifndef CASM_monte_run_management_jsonResultsIO_MI355X
#define CASM_monte_run_management_jsonResultsIO_MI355X

#include <algorithm>
#include <filesystem>
#include <functional>
#include <optional>
#include <set>
#include <string>
#include <vector>

#include "casm/casm_io/json/jsonParser.hh"
#include "casm/monte/checks/io/json/CompletionCheck_json_io.hh"
#include "casm/monte/run_management/RunManager.hh"

namespace CASM {
namespace monte {

inline jsonParser to_json_vector(VectorXd const &v) {
  jsonParser j = jsonParser::array();
  for (Index i = 0; i < v.size(); ++i) j.push_back(v(i));
  return j;
}

/// BasicStatistics.cc:190-206
inline void append_statistics_to_json_arrays(std::optional<BasicStatistics> const &stats,
                                             jsonParser &json) {
  if (!json.contains("mean")) json["mean"] = jsonParser::array();
  if (!json.contains("calculated_precision"))
    json["calculated_precision"] = jsonParser::array();
  if (stats.has_value()) {
    json["mean"].push_back(stats->mean);
    json["calculated_precision"].push_back(stats->calculated_precision);
  } else {
    json["mean"].push_back(jsonParser::null());
    json["calculated_precision"].push_back(jsonParser::null());
  }
}

// --- run_management/Results.hh:165-260 ------------------------------------------
template <typename C, typename S>
bool is_auto_converge_mode(Results<C, S> const &r) {
  return r.completion_check_results.params.requested_precision.size() != 0;
}
template <typename C, typename S>
bool is_requested_to_converge(SamplerComponent const &key, Results<C, S> const &r) {
  auto const &rp = r.completion_check_results.params.requested_precision;
  return rp.find(key) != rp.end();
}
template <typename C, typename S>
Index N_samples(Results<C, S> const &r) {
  return get_n_samples(r.samplers);
}
template <typename C, typename S>
Index N_samples_for_statistics(Results<C, S> const &r) {
  if (is_auto_converge_mode(r))
    return r.completion_check_results.convergence_check_results.N_samples_for_statistics;
  return N_samples(r);
}
template <typename C, typename S>
bool all_equilibrated(Results<C, S> const &r) {
  return r.completion_check_results.equilibration_check_results.all_equilibrated;
}
template <typename C, typename S>
bool all_converged(Results<C, S> const &r) {
  return r.completion_check_results.convergence_check_results.all_converged;
}
template <typename C, typename S>
double acceptance_rate(Results<C, S> const &r) {
  return static_cast<double>(r.n_accept) / static_cast<double>(r.n_accept + r.n_reject);
}

/// Results.hh:261-341: statistics of every component of one sampled quantity
template <typename C, typename S>
struct QuantityStats {
  QuantityStats(std::string const &name, Sampler const &sampler, Results<C, S> const &r)
      : shape(sampler.shape()),
        is_scalar(shape.size() == 0),
        component_names(sampler.component_names()) {
    auto calc = r.completion_check_results.params.calc_statistics_f;
    if (calc == nullptr) throw std::runtime_error("Error in QuantityStats: calc_statistics_f == nullptr");
    static const VectorXd empty_weight;
    auto stats_of_tail = [&](Index comp, Index n) {
      VectorXd tail = sampler.component(comp).tail(n);
      if (r.sample_weight.n_samples() == 0) return calc(tail, empty_weight);
      return calc(tail, r.sample_weight.component(0).tail(n));
    };
    for (Index i = 0; i < (Index)component_names.size(); ++i) {
      SamplerComponent key(name, i, component_names[i]);
      if (is_auto_converge_mode(r)) {
        if (N_samples_for_statistics(r) == 0) {
          is_converged.push_back(is_requested_to_converge(key, r) ? std::optional<bool>(false)
                                                                  : std::nullopt);
          component_stats.push_back(std::nullopt);
        } else if (is_requested_to_converge(key, r)) {
          auto const &cr = r.completion_check_results.convergence_check_results
                               .individual_results.find(key)
                               ->second;
          is_converged.push_back(cr.is_converged);
          component_stats.push_back(cr.stats);
        } else {
          is_converged.push_back(std::nullopt);
          component_stats.push_back(stats_of_tail(i, N_samples_for_statistics(r)));
        }
      } else {
        is_converged.push_back(std::nullopt);
        component_stats.push_back(stats_of_tail(i, sampler.n_samples()));
      }
    }
  }
  std::vector<Index> shape;
  bool is_scalar;
  std::vector<std::string> component_names;
  std::vector<std::optional<bool>> is_converged;
  std::vector<std::optional<S>> component_stats;
};

/// Configuration to JSON for trajectory.json: {"shape": [...], "occupation": [...]}
template <typename ConfigType>
jsonParser config_to_json(ConfigType const &config) {
  jsonParser j;
  j["shape"] = config.shape;
  j["occupation"] = config.occupation();
  return j;
}

template <typename ResultsType>
class jsonResultsIO {
 public:
  typedef ResultsType results_type;

  jsonResultsIO(std::filesystem::path output_dir, bool write_trajectory = false,
                bool write_observations = false)
      : m_output_dir(output_dir),
        m_write_trajectory(write_trajectory),
        m_write_observations(write_observations) {}

  /// jsonResultsIO_impl.hh:340-355
  void write(results_type const &results, ValueMap const &conditions, Index run_index) {
    write_summary(results, conditions);
    if (m_write_trajectory) write_trajectory(results, run_index);
    if (m_write_observations) write_observations(results, run_index);
  }

  /// The hook for SamplingFixtureParams::results_io.
  std::function<void(results_type const &, ValueMap const &, Index)> writer() {
    auto self = *this;
    return [self](results_type const &r, ValueMap const &c, Index i) mutable {
      self.write(r, c, i);
    };
  }

  jsonParser read_summary() const {
    auto path = m_output_dir / "summary.json";
    if (!std::filesystem::exists(path)) return jsonParser::object();
    return jsonParser::read(path);
  }

  /// summary.json, jsonResultsIO_impl.hh:371-453
  void write_summary(results_type const &results, ValueMap const &conditions) {
    jsonParser json = read_summary();
    for (auto key : {"conditions", "statistics", "completion_check_results", "analysis"})
      ensure_object(json, key);
    for (auto const &c : conditions.scalar_values)
      append_condition(c.first, VectorXd{c.second}, {}, {}, json["conditions"]);
    for (auto const &c : conditions.vector_values) {
      std::vector<Index> shape{c.second.size()};
      auto it = results.sampling_functions.find(c.first);
      std::vector<std::string> names = default_component_names(shape);
      if (it != results.sampling_functions.end()) {
        if ((Index)it->second.component_names.size() != c.second.size())
          throw std::runtime_error("Error in get_vector_component_names: Dimension of \"" +
                                   c.first + "\" does not match the corresponding sampling function.");
        names = it->second.component_names;
      }
      append_condition(c.first, c.second, shape, names, json["conditions"]);
    }
    for (auto const &q : results.samplers) append_statistics(q.first, *q.second, results, json["statistics"]);
    append_completion_check_results(results, json["completion_check_results"]);
    append_results_analysis(results, json["analysis"]);
    (json).write(m_output_dir / "summary.json", -1);
  }

  /// run.<i>/trajectory.json, jsonResultsIO_impl.hh:457-462
  void write_trajectory(results_type const &results, Index run_index) {
    jsonParser json = jsonParser::array();
    for (auto const &config : results.sample_trajectory) json.push_back(config_to_json(config));
    json.write(run_dir(run_index) / "trajectory.json", -1);
  }

  /// run.<i>/observations.json, jsonResultsIO_impl.hh:464-526
  void write_observations(results_type const &results, Index run_index) {
    jsonParser json = jsonParser::object();
    if (results.sample_count.size()) json["count"] = results.sample_count;
    if (results.sample_time.size()) json["time"] = results.sample_time;
    if (results.sample_weight.n_samples()) json["weight"] = to_json_vector(results.sample_weight.component(0));
    if (results.sample_clocktime.size()) json["clocktime"] = results.sample_clocktime;
    for (auto const &p : results.samplers) {
      Sampler const &s = *p.second;
      json[p.first]["shape"] = s.shape();
      if (s.shape().size() == 0) {
        json[p.first]["value"] = to_json_vector(s.component(0));
      } else {
        json[p.first]["component_names"] = s.component_names();
        jsonParser rows = jsonParser::array();  // one row per sample
        for (CountType k = 0; k < s.n_samples(); ++k) rows.push_back(to_json_vector(s.sample(k)));
        json[p.first]["value"] = rows;
      }
    }
    for (auto const &p : results.json_samplers) {
      jsonParser vals = jsonParser::array();
      for (auto const &text : p.second->values) vals.push_back(jsonParser::parse(text));
      json[p.first]["value"] = vals;
    }
    json.write(run_dir(run_index) / "observations.json", -1);
  }

  std::filesystem::path run_dir(Index run_index) const {
    return m_output_dir / ("run." + std::to_string(run_index));
  }

 private:
  std::filesystem::path m_output_dir;
  bool m_write_trajectory;
  bool m_write_observations;

  static void ensure_object(jsonParser &json, std::string const &key) {
    if (!json.contains(key))
      json[key].put_obj();
    else if (!json[key].is_obj())
      throw std::runtime_error("JSON Error: \"" + key + "\" is expected to be an object.");
  }
  static void ensure_array(jsonParser &json, std::string const &key) {
    if (!json.contains(key))
      json[key].put_array();
    else if (!json[key].is_array())
      throw std::runtime_error("JSON Error: \"" + key + "\" is expected to be an array.");
  }

  /// jsonResultsIO_impl.hh:46-100
  static void append_condition(std::string const &name, VectorXd const &value,
                               std::vector<Index> const &shape,
                               std::vector<std::string> const &names, jsonParser &json) {
    ensure_object(json, name);
    auto &j = json[name];
    j["shape"] = shape;
    if (shape.size() == 0) {
      ensure_array(j, "value");
      j["value"].push_back(value(0));
    } else {
      j["component_names"] = names;
      for (Index i = 0; i < (Index)names.size(); ++i) {
        ensure_array(j, names[i]);
        j[names[i]].push_back(value(i));
      }
    }
  }

  /// jsonResultsIO_impl.hh:135-196
  static void append_statistics(std::string const &name, Sampler const &sampler,
                                results_type const &results, jsonParser &json) {
    QuantityStats<typename results_type::config_type, typename results_type::stats_type> q(
        name, sampler, results);
    ensure_object(json, name);
    auto &qj = json[name];
    qj["shape"] = q.shape;
    auto append = [&](jsonParser &t, Index i) {
      append_statistics_to_json_arrays(q.component_stats[i], t);
      if (q.is_converged[i].has_value()) {
        ensure_array(t, "is_converged");
        t["is_converged"].push_back(*q.is_converged[i]);
      }
    };
    if (q.is_scalar) {
      ensure_object(qj, "value");
      append(qj["value"], 0);
    } else {
      qj["component_names"] = q.component_names;
      for (Index i = 0; i < (Index)q.component_names.size(); ++i) {
        ensure_object(qj, q.component_names[i]);
        append(qj[q.component_names[i]], i);
      }
    }
  }

  /// jsonResultsIO_impl.hh:198-258
  static void append_completion_check_results(results_type const &r, jsonParser &json) {
    if (is_auto_converge_mode(r)) {
      for (auto key : {"all_equilibrated", "all_converged", "N_samples_for_all_to_equilibrate"})
        ensure_array(json, key);
      json["all_equilibrated"].push_back(all_equilibrated(r));
      json["all_converged"].push_back(all_converged(r));
      if (all_equilibrated(r))
        json["N_samples_for_all_to_equilibrate"].push_back(
            r.completion_check_results.equilibration_check_results
                .N_samples_for_all_to_equilibrate);
      else
        json["N_samples_for_all_to_equilibrate"].push_back("did_not_equilibrate");
    }
    for (auto key : {"N_samples", "N_samples_for_statistics", "acceptance_rate",
                     "initial_memory_used_MiB", "final_memory_used_MiB", "elapsed_clocktime",
                     "count"})
      ensure_array(json, key);
    json["N_samples"].push_back(N_samples(r));
    json["N_samples_for_statistics"].push_back(N_samples_for_statistics(r));
    json["acceptance_rate"].push_back(acceptance_rate(r));
    json["initial_memory_used_MiB"].push_back(r.initial_memory_used_MiB);
    json["final_memory_used_MiB"].push_back(r.final_memory_used_MiB);
    json["elapsed_clocktime"].push_back(jsonParser(r.elapsed_clocktime));
    json["count"].push_back(r.sample_count.size() ? jsonParser(r.sample_count.back())
                                                  : jsonParser::null());
    if (r.sample_time.size()) {
      ensure_array(json, "time");
      json["time"].push_back(r.sample_time.back());
    }
  }

  /// jsonResultsIO_impl.hh:260-330
  static void append_results_analysis(results_type const &r, jsonParser &json) {
    for (auto const &p : r.analysis) {
      auto it = r.analysis_functions.find(p.first);
      if (it == r.analysis_functions.end())
        throw std::runtime_error(
            "Error in append_results_analysis_to_json: No matching analysis function found for '" +
            p.first + "'.");
      ensure_object(json, p.first);
      jsonParser &vj = json[p.first];
      vj["shape"] = it->second.shape;
      const bool dne = is_auto_converge_mode(r) && !all_equilibrated(r);
      if (it->second.shape.size() == 0) {
        ensure_array(vj, "value");
        vj["value"].push_back(dne ? jsonParser("did_not_equilibrate") : jsonParser(p.second(0)));
      } else {
        vj["component_names"] = it->second.component_names;
        for (Index i = 0; i < (Index)it->second.component_names.size(); ++i) {
          auto const &cn = it->second.component_names[i];
          ensure_array(vj, cn);
          vj[cn].push_back(dne ? jsonParser("did_not_equilibrate") : jsonParser(p.second(i)));
        }
      }
    }
  }
};

}  // namespace monte
}  // namespace CASM

#endif
