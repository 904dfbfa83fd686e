"""Writes the golden fixtures of tests/golden/ from the oracle (oracle/refcpu.hh,
which runs the reference loop over the host libstdc++ <random> and glibc exp).

  mt19937_64.json : known outputs of std::mt19937_64 for seeds 42, 5489 and
                    1000..1063 (first 8 outputs, the 10000th, and a SHA-256 of
                    the first 10^4 outputs as little-endian bytes);
  c1.json         : the 25x25 reference run (J=0.1 eV, mu=0, T=2000 K, all +1,
                    seed 42, 11000 passes sampled every pass), both energy
                    variants: counts, final engine words, hashes of the
                    integer and FP64 series and of the final lattice, the first
                    and last 50 sampled values as hex floats, and the first 200
                    attempts (site, old spin, neighbour sum, drew, accepted);
  c2.json         : the 64-chain 128x128 scan (T = 500 + a*2500/7,
                    mu = -0.4 + b*0.8/7, seed 1000 + 8a + b, all +1) after 2000
                    passes: per chain counts, hashes of the (S, M) series, of
                    the potential_energy series and of the final lattice and
                    engine words.

Run from the repository root: python tests/golden/make_golden.py
(about two minutes on 8 cores; records the host compiler and libstdc++ used).
"""
import concurrent.futures as cf
import hashlib
import json
import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def hexs(a):
    return [float(x).hex() for x in a]


def provenance():
    try:
        gxx = subprocess.run(["g++", "--version"], stdout=subprocess.PIPE, text=True).stdout
        gxx = gxx.splitlines()[0]
    except Exception:
        gxx = "unknown"
    return {"compiler": gxx, "platform": platform.platform(), "KB": O.kb(),
            "generator": "tests/golden/make_golden.py"}


def c2_conditions():
    out = []
    for a in range(8):
        for b in range(8):
            out.append(dict(T=500.0 + a * 2500.0 / 7.0, mu=-0.4 + b * 0.8 / 7.0,
                            seed=1000 + 8 * a + b))
    return out


def c2_entry(c, passes=2000):
    o = O.sgc_run(128, 128, J=0.1, use_nlist=True, temperature=c["T"], mu=c["mu"],
                  seed=c["seed"], max_count=passes, capacity=passes + 4)
    sm = np.stack([o["S"], o["M"]], axis=1).astype(np.int64)
    return dict(T=c["T"], mu=c["mu"], seed=c["seed"], n_pass=o["n_pass"],
                n_accept=o["n_accept"], n_reject=o["n_reject"], SM_sha256=sha(sm),
                potential_energy_sha256=sha(o["potential_energy"]),
                potential_energy_last=o["potential_energy"][-1].hex(),
                final_occ_sha256=sha(o["final_occ"].astype(np.int8)),
                engine_sha256=sha(o["engine"]))


def main():
    seeds = [42, 5489] + list(range(1000, 1064))
    mt = {}
    for s in seeds:
        x = O.mt64_outputs(s, 10000)
        mt[str(s)] = dict(first8=[int(v) for v in x[:8]], out10000=int(x[-1]),
                          sha256_first10000=sha(x.astype("<u8")))
    with open(os.path.join(HERE, "mt19937_64.json"), "w") as f:
        json.dump(dict(provenance=provenance(), seeds=mt), f, indent=1)

    c1 = {}
    for use_nlist in (True, False):
        o = O.sgc_run(25, 25, J=0.1, use_nlist=use_nlist, temperature=2000.0, mu=0.0, seed=42,
                      max_count=11000, trace_len=200)
        sm = np.stack([o["S"], o["M"]], axis=1).astype(np.int64)
        c1["nlist" if use_nlist else "fold"] = dict(
            n_pass=o["n_pass"], n_samples=o["n_samples"], n_accept=o["n_accept"],
            n_reject=o["n_reject"], engine=[int(v) for v in o["engine"]],
            SM_sha256=sha(sm), final_occ_sha256=sha(o["final_occ"].astype(np.int8)),
            **{f"{k}_sha256": sha(o[k]) for k in
               ("formation_energy", "param_composition", "potential_energy")},
            **{f"{k}_first50": hexs(o[k][:50]) for k in
               ("formation_energy", "param_composition", "potential_energy")},
            **{f"{k}_last50": hexs(o[k][-50:]) for k in
               ("formation_energy", "param_composition", "potential_energy")},
            trace200=o["trace"].tolist())
    with open(os.path.join(HERE, "c1.json"), "w") as f:
        json.dump(dict(provenance=provenance(), config="25x25 J=0.1 mu=0 T=2000 seed 42 "
                       "11000 passes", runs=c1), f, indent=1)

    with cf.ThreadPoolExecutor(os.cpu_count() or 4) as ex:
        chains = list(ex.map(c2_entry, c2_conditions()))
    with open(os.path.join(HERE, "c2.json"), "w") as f:
        json.dump(dict(provenance=provenance(), config="64 chains 128x128 J=0.1 all +1, "
                       "2000 passes, sampled every pass", chains=chains), f, indent=1)
    print("wrote", [n for n in os.listdir(HERE) if n.endswith(".json")])


if __name__ == "__main__":
    main()
