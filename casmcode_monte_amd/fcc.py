"""Binary FCC cluster expansion: its default ECIs and the 256-chain batch of
uneven supercells.

The model (include/casm_mi355x/mcgpu.h, MCG_LATTICE_FCC) is build-defined:
libcasm-monte itself ships only the 2D Ising model.  ECIs: empty V0 = 0.0 eV,
point V1 = -0.1 eV, and the nearest/second/third-neighbour pair and
nearest-neighbour triplet ECIs drawn from N(0, 0.05 eV) by
numpy.random.default_rng(7).normal (kept here as exact literals; the fixture
tests/golden/fcc_eci.json and its generator check them).
"""
import numpy as np

ECI_NAMES = ("V0", "V1", "V_nn", "V_2nn", "V_3nn", "V_triplet")
DEFAULT_ECI = (
    0.0,
    -0.1,
    float.fromhex("0x1.01fb5b01996adp-14"),
    float.fromhex("0x1.e976f5d5eb397p-7"),
    float.fromhex("-0x1.c125c0159bbc0p-7"),
    float.fromhex("-0x1.6cc952a279a68p-5"),
)


def draw_eci(seed=7):
    """The default ECIs regenerated: V0 = 0, V1 = -0.1, then 4 draws of N(0, 0.05)."""
    v = np.random.default_rng(seed).normal(0.0, 0.05, 4)
    return (0.0, -0.1) + tuple(float(x) for x in v)


def batch_conditions(n=256, seed=8):
    """The 256-chain batch: chain i has L_i = 16 + (i mod 17) conventional cells per
    edge, T ~ U(300, 2000) K and mu ~ U(-1, 1) eV from default_rng(seed), Philox
    key 0xFCC000 + i, all sites empty at the start."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(300.0, 2000.0, n)
    mu = rng.uniform(-1.0, 1.0, n)
    out = []
    for i in range(n):
        L = 16 + (i % 17)
        out.append(dict(shape=(L, L, L), temperature=float(T[i]),
                        exchange_potential=float(mu[i]), seed=0xFCC000 + i))
    return out
