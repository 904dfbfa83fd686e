"""Runs the C++ drop-in test program (tests/cpp/test_dropin.cpp, built by
__graft_entry__.build()): the reference-named classes of include/casm/monte/
against the oracle."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "dropin_tests")


def run(flag):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    p = subprocess.run([BIN, flag], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=600)
    print(p.stdout)
    assert p.returncode == 0 and "ALL OK" in p.stdout, p.stdout


def test_dropin_host_logic_matches_oracle():
    """CompletionCheck, BasicStatistics, equilibration check: same decisions and
    same bits as the oracle on synthetic series; the max_count=12 known answer."""
    run("--cpu")


@pytest.mark.gpu
def test_dropin_calculator_on_gpu_matches_oracle():
    """SemiGrandCanonicalCalculator::run and occupation_metropolis with a
    RunManager on the MI355X reproduce the reference loop bit for bit."""
    run("--gpu")
