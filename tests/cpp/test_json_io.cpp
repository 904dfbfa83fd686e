// json_io_tests <dir>: writes two runs of a sampling fixture's results with
// jsonResultsIO into <dir> (summary.json appended per run, run.<i>/observations.json
// and trajectory.json), reads a CompletionCheckParams from JSON and checks a
// jsonParser round trip.  The file contents are checked by tests/test_json_io.py.
#include <cassert>
#include <iostream>

#include "casm/monte/ising_cpp/model.hh"
#include "casm/monte/run_management/io/json/jsonResultsIO.hh"

using namespace CASM;
using namespace CASM::monte;
using namespace CASM::monte::ising_cpp;

typedef Results<IsingConfiguration, BasicStatistics> results_type;

static void check(bool ok, const char *what) {
  if (!ok) {
    std::cout << "FAILED: " << what << std::endl;
    std::exit(1);
  }
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::filesystem::path dir(argv[1]);

  // jsonParser round trip
  {
    jsonParser j;
    j["a"] = 1.5;
    j["b"] = std::vector<long>{1, 2, 3};
    j["c"]["d"] = std::string("x\"y");
    j["e"] = true;
    j["f"] = jsonParser::null();
    j["g"] = std::vector<double>{0.1, -2e-300, 12345678901234.5};
    std::stringstream ss;
    j.print(ss, 2);
    jsonParser k = jsonParser::parse(ss.str());
    check(k["a"].get<double>() == 1.5, "double");
    check(k["b"].get<std::vector<long>>() == std::vector<long>({1, 2, 3}), "int array");
    check(k["c"]["d"].get<std::string>() == "x\"y", "escaped string");
    check(k["e"].get<bool>() && k["f"].is_null(), "bool/null");
    check(k["g"].get<std::vector<double>>() == std::vector<double>({0.1, -2e-300, 12345678901234.5}),
          "doubles round trip");
  }

  // completion check params from JSON (CompletionCheck_json_io.hh:295-369)
  StateSamplingFunctionMap f;
  f.emplace("param_composition",
            StateSamplingFunction("param_composition", "x", {1}, []() { return VectorXd{0.0}; }));
  f.emplace("potential_energy",
            StateSamplingFunction("potential_energy", "e", {}, []() { return VectorXd{0.0}; }));
  auto ccp = completion_check_params_from_json(
      jsonParser::parse(R"({"cutoff": {"count": {"min": 10, "max": 5000}, "sample": {"min": 100}},
                            "convergence": [{"quantity": "potential_energy", "precision": 0.001},
                                            {"quantity": "param_composition", "precision": 0.002}],
                            "spacing": "linear", "begin": 100, "period": 10})"),
      f);
  check(ccp.cutoff_params.min_count == 10 && ccp.cutoff_params.max_count == 5000, "cutoff count");
  check(ccp.cutoff_params.min_sample == 100 && !ccp.cutoff_params.max_sample, "cutoff sample");
  check(ccp.check_begin == 100 && ccp.check_period == 10 && !ccp.log_spacing, "spacing");
  check(ccp.requested_precision.size() == 2, "requested precision size");
  check(ccp.requested_precision.at(SamplerComponent("param_composition", 0, "0")).abs_precision ==
            0.002,
        "param_composition precision");

  // two runs of results
  jsonResultsIO<results_type> io(dir, true, true);
  for (int run = 0; run < 2; ++run) {
    std::vector<std::string> names{"param_composition", "potential_energy"};
    results_type r(names, f, {}, jsonStateSamplingFunctionMap(),
                   ResultsAnalysisFunctionMap<IsingConfiguration, BasicStatistics>());
    r.reset();
    IsingConfiguration config({4, 4}, 1);
    for (int k = 0; k < 300; ++k) {
      double x = 0.5 + 0.1 * std::sin(0.37 * k + run), e = -0.2 + 0.01 * std::cos(0.11 * k);
      r.samplers.at("param_composition")->push_back(VectorXd{x});
      r.samplers.at("potential_energy")->push_back(e);
      r.sample_count.push_back(k + 1);
      r.sample_clocktime.push_back(0.001 * k);
      if (k % 100 == 0) {
        config.set_occ(k % 16, -1);
        r.sample_trajectory.push_back(config);
      }
    }
    // as left by a completion check with the requested precision of ccp
    r.completion_check_results.params = ccp;
    CompletionCheck<BasicStatistics> cc(ccp);
    r.completion_check_results.equilibration_check_results = equilibration_check(
        default_equilibration_check, ccp.requested_precision, r.samplers, r.sample_weight, true);
    r.completion_check_results.convergence_check_results = convergence_check(
        r.samplers, r.sample_weight, ccp.requested_precision,
        r.completion_check_results.equilibration_check_results.N_samples_for_all_to_equilibrate,
        ccp.calc_statistics_f);
    r.n_accept = 1000 + run;
    r.n_reject = 3000;
    r.elapsed_clocktime = 1.25;
    ValueMap cond;
    cond.scalar_values["temperature"] = 1000.0 + 500.0 * run;
    cond.vector_values["exchange_potential"] = VectorXd{0.1 * run};
    io.write(r, cond, run);
  }
  std::cout << "ALL OK" << std::endl;
  return 0;
}
