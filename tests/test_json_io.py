"""Results files in the reference's formats (jsonResultsIO_impl.hh:360-526) and
CompletionCheckParams read from JSON (CompletionCheck_json_io.hh:295-369),
written by the C++ drop-in (tests/cpp/test_json_io.cpp) and checked here:
appended summary arrays, statistics of the post-equilibration tail equal to
the oracle's BasicStatistics of the same tail, observations and trajectories
as sampled."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "bin", "json_io_tests")


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail("tests/bin/json_io_tests not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("results")
    p = subprocess.run([BIN, str(d)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "ALL OK" in p.stdout, p.stdout + p.stderr
    return d


def series(run):
    k = np.arange(300)
    x = 0.5 + 0.1 * np.sin(0.37 * k + run)
    e = -0.2 + 0.01 * np.cos(0.11 * k)
    return x, e


def test_summary(out):
    s = json.load(open(out / "summary.json"))
    assert set(s) == {"conditions", "statistics", "completion_check_results", "analysis"}
    assert s["conditions"]["temperature"] == {"shape": [], "value": [1000.0, 1500.0]}
    ep = s["conditions"]["exchange_potential"]
    assert ep["shape"] == [1] and ep["component_names"] == ["0"] and ep["0"] == [0.0, 0.1]
    cc = s["completion_check_results"]
    assert cc["N_samples"] == [300, 300] and cc["count"] == [300, 300]
    assert cc["acceptance_rate"] == [1000 / 4000, 1001 / 4001]
    assert cc["elapsed_clocktime"] == [1.25, 1.25]
    n_stat = cc["N_samples_for_statistics"]
    for run in (0, 1):
        x, e = series(run)
        assert cc["N_samples_for_all_to_equilibrate"][run] == 300 - n_stat[run]
        for name, vals, node in (("param_composition", x, s["statistics"]["param_composition"]["0"]),
                                 ("potential_energy", e, s["statistics"]["potential_energy"]["value"])):
            tail = vals[300 - n_stat[run]:]
            m, prec = O.basic_statistics(tail)
            assert math.isclose(node["mean"][run], m, rel_tol=1e-13, abs_tol=1e-15)
            assert math.isclose(node["calculated_precision"][run], prec, rel_tol=1e-12)
            req = 0.002 if name == "param_composition" else 0.001
            assert node["is_converged"][run] == (prec < req)
    assert s["statistics"]["param_composition"]["shape"] == [1]
    assert s["statistics"]["potential_energy"]["shape"] == []


def test_observations_and_trajectory(out):
    for run in (0, 1):
        o = json.load(open(out / f"run.{run}" / "observations.json"))
        x, e = series(run)
        assert o["count"] == list(range(1, 301))
        assert o["param_composition"]["shape"] == [1]
        assert o["param_composition"]["component_names"] == ["0"]
        np.testing.assert_array_equal(np.array(o["param_composition"]["value"])[:, 0], x)
        assert o["potential_energy"]["shape"] == []
        np.testing.assert_array_equal(np.array(o["potential_energy"]["value"]), e)
        t = json.load(open(out / f"run.{run}" / "trajectory.json"))
        assert len(t) == 3 and all(c["shape"] == [4, 4] for c in t)
        assert sum(v == -1 for v in t[-1]["occupation"]) == 3


def test_completion_check_params_round_trip(out):
    """CompletionCheckParams to_json (CompletionCheck_json_io.hh:371-417) of the
    parameters read from JSON above."""
    p = json.load(open(out / "completion_check_params.json"))
    assert p["cutoff"] == {"count": {"min": 10, "max": 5000}, "sample": {"min": 100}}
    assert p["spacing"] == "linear" and p["begin"] == 100 and p["period"] == 10
    conv = {c["quantity"]: c for c in p["convergence"]}
    assert conv["potential_energy"]["abs_precision"] == 0.001
    assert conv["param_composition"]["abs_precision"] == 0.002
    assert conv["param_composition"]["component_index"] == [0]
    assert conv["param_composition"]["component_name"] == ["0"]


def test_completion_check_results_json(out):
    """CompletionCheckResults to_json (CompletionCheck_json_io.hh:420-440, with the
    equilibration and convergence results of EquilibrationCheck_json_io.cc and
    ConvergenceCheck_json_io.hh:33-62); the statistics equal the oracle's
    BasicStatistics of the same post-equilibration tail."""
    r = json.load(open(out / "completion_check_results.json"))
    for key in ("has_all_minimums_met", "has_any_maximum_met", "count", "time", "clocktime",
                "n_samples", "is_complete", "n_samples_at_convergence_check",
                "equilibration_check_results", "convergence_check_results"):
        assert key in r, key
    assert r["count"] == 400 and r["n_samples"] == 400 and r["time"] is None
    eq = r["equilibration_check_results"]
    assert eq["all_equilibrated"] is True
    n_eq = eq["N_samples_for_all_to_equilibrate"]
    series = {
        "param_composition": np.array([0.5 + 0.1 * math.sin(0.37 * k) for k in range(400)]),
        "potential_energy": np.array([-0.2 + 0.02 * math.cos(0.9 * k) for k in range(400)]),
    }
    cv = r["convergence_check_results"]
    assert cv["N_samples_for_statistics"] == 400 - n_eq
    for ind in cv["individual_results"]:
        mean, prec = O.basic_statistics(series[ind["sampler_name"]][n_eq:], 0.95)
        assert ind["stats"]["mean"] == pytest.approx(mean, rel=1e-13)
        assert ind["stats"]["calculated_precision"] == pytest.approx(prec, rel=1e-10)
        assert ind["is_converged"] == (prec < ind["requested_precision"]["abs_precision"])


def test_sampling_fixture_params_json(out):
    """SamplingFixtureParams read from JSON (SamplingFixtureParams_json_io.hh:37-147)
    and written back; its results_io writes a run into its output_dir."""
    p = json.load(open(out / "fixture_params.json"))
    assert p["sampling"] == {"sample_by": "pass", "spacing": "linear", "begin": 0, "period": 1,
                             "quantities": ["param_composition"]}
    assert p["completion_check"]["cutoff"] == {"count": {"max": 200}}
    assert p["results_io"] == {"method": "json"}
    assert p["log"]["frequency_in_s"] == 0.0 and p["log"]["file"].endswith("status.json")
    s = json.load(open(out / "fixture_out" / "summary.json"))
    assert s["conditions"]["temperature"]["value"] == [700.0]
    assert os.path.exists(out / "fixture_out" / "run.0" / "observations.json")
