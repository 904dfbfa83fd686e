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
