"""Full-size EXACT-mode parity against the golden fixtures (no oracle run on the
GPU host): the 25x25 reference run (both energy variants, 11000 passes) and
all 64 chains of the 128x128 scan at their full 2000 passes, in one launch."""
import hashlib
import json
import os

import numpy as np
import pytest

import casmcode_monte_amd as cm

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(G, name)) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("variant", ["nlist", "fold"])
def test_c1_full_golden(variant):
    e = load("c1.json")["runs"][variant]
    print("case: golden C1", variant, flush=True)
    r = cm.Run([dict(shape=(25, 25), temperature=2000.0, exchange_potential=0.0, seed=42)],
               mode=cm.MODE_EXACT, J1=0.1, use_nlist=(variant == "nlist"))
    r.run(11000)
    rec = r.samples(0)
    assert r.counters(0) == (e["n_pass"], e["n_accept"], e["n_reject"])
    assert [int(v) for v in r.rng_dump(0)] == e["engine"]
    assert sha(np.stack([rec["S1"], rec["M"]], axis=1).astype(np.int64)) == e["SM_sha256"]
    obs = r.observables(0, rec)
    for k in ("formation_energy", "param_composition", "potential_energy"):
        assert sha(obs[k]) == e[f"{k}_sha256"], k
        assert [float(v).hex() for v in obs[k][-50:]] == e[f"{k}_last50"]
    assert sha(r.get_occupation(0)) == e["final_occ_sha256"]


def test_c2_full_golden_64_chains():
    chains = load("c2.json")["chains"]
    print("case: golden C2 64 x 128^2 x 2000", flush=True)
    r = cm.Run([dict(shape=(128, 128), temperature=c["T"], exchange_potential=c["mu"],
                     seed=c["seed"]) for c in chains], mode=cm.MODE_EXACT, J1=0.1)
    r.run(2000)
    for i, c in enumerate(chains):
        rec = r.samples(i)
        _, acc, rej = r.counters(i)
        assert (acc, rej) == (c["n_accept"], c["n_reject"]), i
        assert sha(np.stack([rec["S1"], rec["M"]], axis=1).astype(np.int64)) == c["SM_sha256"], i
        assert sha(r.observables(i, rec)["potential_energy"]) == c["potential_energy_sha256"], i
        assert sha(r.get_occupation(i)) == c["final_occ_sha256"], i
        assert sha(r.rng_dump(i)) == c["engine_sha256"], i
