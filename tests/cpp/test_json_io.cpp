// json_io_tests <dir>: writes two runs of a sampling fixture's results with
// jsonResultsIO into <dir> (summary.json appended per run, run.<i>/observations.json
// and trajectory.json), reads a CompletionCheckParams, SamplingParams and
// SamplingFixtureParams from JSON, writes the completion check parameters and
// results as JSON, and checks a jsonParser round trip.  The file contents are
// checked by tests/test_json_io.py.
#include <cassert>
#include <iostream>

#include "casm/monte/ising_cpp/model.hh"
#include "casm/monte/run_management/io/json/SamplingFixtureParams_json_io.hh"
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

  // SamplingParams from JSON (SamplingParams_json_io.cc:70-176) and back
  {
    std::set<std::string> names{"param_composition", "potential_energy"};
    SamplingParams sp = sampling_params_from_json(
        jsonParser::parse(R"({"sample_by": "pass", "spacing": "log", "period": 10.0,
                              "quantities": ["potential_energy"]})"),
        names, {}, false);
    check(sp.sample_mode == SAMPLE_MODE::BY_PASS && sp.sample_method == SAMPLE_METHOD::LOG,
          "sampling mode/method");
    check(sp.begin == 0.0 && sp.period == 10.0 && sp.shift == 10.0 &&
              sp.base == std::pow(10.0, 0.1),
          "sampling defaults");
    check(sp.sampler_names == std::vector<std::string>{"potential_energy"}, "quantities");
    auto fails = [&](const char *text, const char *needle) {
      try {
        sampling_params_from_json(jsonParser::parse(text), names, {}, false);
      } catch (std::runtime_error const &e) {
        return std::string(e.what()).find(needle) != std::string::npos;
      }
      return false;
    };
    check(fails(R"({"sample_by": "pass", "period": 1, "quantities": ["nope"]})",
                "\"nope\" is not a sampling option"),
          "unknown quantity");
    check(fails(R"({"sample_by": "time", "period": 1})", "must be one of \"pass\" or \"step\""),
          "time sampling not allowed");
    check(fails(R"({"sample_by": "pass", "spacing": "log", "period": 1.0})", "must > 1.0"),
          "log period");
    check(fails(R"({"sample_by": "pass"})", "period"), "period required");
    jsonParser back;
    to_json(sp, back);
    check(back["spacing"].get<std::string>() == "log" && back["sample_by"].get<std::string>() == "pass",
          "SamplingParams to_json");
  }

  // SamplingFixtureParams from JSON (SamplingFixtureParams_json_io.hh:37-147)
  {
    jsonParser fj = jsonParser::parse(R"({
      "sampling": {"sample_by": "pass", "period": 1, "quantities": ["param_composition"]},
      "completion_check": {"cutoff": {"count": {"max": 200}},
                           "convergence": [{"quantity": "param_composition", "abs_precision": 0.01}]},
      "analysis": {"functions": []},
      "results_io": {"method": "json", "kwargs": {"output_dir": "OUT", "write_observations": true}},
      "log": {"file": "STATUS", "frequency_in_s": 0.0}})");
    fj["results_io"]["kwargs"]["output_dir"] = (dir / "fixture_out").string();
    fj["log"]["file"] = (dir / "fixture_out" / "status.json").string();
    auto params = sampling_fixture_params_from_json<IsingConfiguration>(
        fj, "thermo", f, jsonStateSamplingFunctionMap(),
        ResultsAnalysisFunctionMap<IsingConfiguration, BasicStatistics>(), false);
    check(params.label == "thermo" && params.sampling_params.sampler_names.size() == 1,
          "fixture sampling params");
    check(params.completion_check_params.cutoff_params.max_count == 200, "fixture cutoff");
    check(params.results_io != nullptr, "fixture results_io");
    check(params.method_log.log_frequency == 0.0, "fixture log frequency");
    jsonParser pj;
    to_json(params, pj);
    pj.write(dir / "fixture_params.json");
    // the results writer the fixture was given: one run into output_dir
    std::vector<std::string> names{"param_composition"};
    results_type r(names, f, {}, jsonStateSamplingFunctionMap(),
                   ResultsAnalysisFunctionMap<IsingConfiguration, BasicStatistics>());
    r.reset();
    for (int k = 0; k < 50; ++k) {
      r.samplers.at("param_composition")->push_back(VectorXd{0.25 + 0.001 * (k % 7)});
      r.sample_count.push_back(k + 1);
    }
    r.completion_check_results.params = params.completion_check_params;
    ValueMap cond;
    cond.scalar_values["temperature"] = 700.0;
    params.results_io(r, cond, 0);
    bool threw = false;
    try {
      jsonParser bad = fj;
      bad["results_io"]["method"] = "hdf5";
      sampling_fixture_params_from_json<IsingConfiguration>(
          bad, "x", f, jsonStateSamplingFunctionMap(),
          ResultsAnalysisFunctionMap<IsingConfiguration, BasicStatistics>(), false);
    } catch (std::runtime_error const &e) {
      threw = std::string(e.what()).find("must be one of: \"json\"") != std::string::npos;
    }
    check(threw, "unknown results_io method");
  }

  // completion check parameters and results as JSON (CompletionCheck_json_io.hh:371-440)
  {
    jsonParser pj;
    to_json(ccp, pj);
    pj.write(dir / "completion_check_params.json");
    CompletionCheck<BasicStatistics> cc(ccp);
    SamplerMap samplers;
    samplers.emplace("param_composition", std::make_shared<Sampler>(std::vector<Index>{1}));
    samplers.emplace("potential_energy", std::make_shared<Sampler>(std::vector<Index>{}));
    for (int k = 0; k < 400; ++k) {
      samplers.at("param_composition")->push_back(VectorXd{0.5 + 0.1 * std::sin(0.37 * k)});
      samplers.at("potential_energy")->push_back(VectorXd{-0.2 + 0.02 * std::cos(0.9 * k)});
    }
    Sampler weight(std::vector<Index>{});
    Log log;
    cc.is_complete(samplers, weight, CountType(400), log);
    jsonParser rj;
    to_json(cc.results(), rj);
    rj.write(dir / "completion_check_results.json");
  }
  std::cout << "ALL OK" << std::endl;
  return 0;
}
