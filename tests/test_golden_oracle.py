"""The oracle reproduces the committed golden fixtures (tests/golden/*.json,
written by tests/golden/make_golden.py): pins the restatement against drift."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(G, name)) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_mt19937_64_fixture():
    d = load("mt19937_64.json")["seeds"]
    assert d["5489"]["out10000"] == 9981545732273789042  # C++ standard known answer
    for s, e in d.items():
        x = O.mt64_outputs(int(s), 10000)
        assert [int(v) for v in x[:8]] == e["first8"]
        assert sha(x.astype("<u8")) == e["sha256_first10000"]


@pytest.mark.parametrize("variant", ["nlist", "fold"])
def test_c1_fixture(variant):
    e = load("c1.json")["runs"][variant]
    o = O.sgc_run(25, 25, J=0.1, use_nlist=(variant == "nlist"), temperature=2000.0, mu=0.0,
                  seed=42, max_count=11000, trace_len=200)
    assert (o["n_pass"], o["n_accept"], o["n_reject"]) == (e["n_pass"], e["n_accept"],
                                                           e["n_reject"])
    assert [int(v) for v in o["engine"]] == e["engine"]
    assert sha(np.stack([o["S"], o["M"]], axis=1).astype(np.int64)) == e["SM_sha256"]
    assert o["trace"].tolist() == e["trace200"]
    for k in ("formation_energy", "param_composition", "potential_energy"):
        assert sha(o[k]) == e[f"{k}_sha256"]
        assert [float(v).hex() for v in o[k][:50]] == e[f"{k}_first50"]


@pytest.mark.parametrize("idx", [0, 37])
def test_c2_fixture_chains(idx):
    c = load("c2.json")["chains"][idx]
    o = O.sgc_run(128, 128, J=0.1, use_nlist=True, temperature=c["T"], mu=c["mu"],
                  seed=c["seed"], max_count=2000, capacity=2004)
    assert (o["n_accept"], o["n_reject"]) == (c["n_accept"], c["n_reject"])
    assert sha(np.stack([o["S"], o["M"]], axis=1).astype(np.int64)) == c["SM_sha256"]
    assert sha(o["final_occ"].astype(np.int8)) == c["final_occ_sha256"]
    assert sha(o["engine"]) == c["engine_sha256"]
