"""Writes tests/golden/fcc_eci.json: the FCC cluster-expansion ECIs (V0 = 0,
V1 = -0.1 eV, four draws of N(0, 0.05 eV) from numpy.random.default_rng(7))
as decimal and hex floats, with the numpy version that drew them.

Run from the repository root: python tests/golden/make_fcc_eci.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

v = np.random.default_rng(7).normal(0.0, 0.05, 4)
eci = [0.0, -0.1] + [float(x) for x in v]
out = dict(names=["V0", "V1", "V_nn", "V_2nn", "V_3nn", "V_triplet"], eci=eci,
           eci_hex=[x.hex() for x in eci], numpy=np.__version__,
           generator="numpy.random.default_rng(7).normal(0.0, 0.05, 4)")
with open(os.path.join(HERE, "fcc_eci.json"), "w") as f:
    json.dump(out, f, indent=1)
print(json.dumps(out))
