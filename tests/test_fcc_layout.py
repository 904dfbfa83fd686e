"""CPU checks of the FCC kernel's work decomposition (a Python emulation of the
index arithmetic of casmcode_monte_amd/csrc/fcc.hip):

  - the ring colouring (six colours, the last two on the seam of a ring of odd
    L) puts equal colours 0 or >= 3 apart on every ring 2L >= 6;
  - no two sites of one colour class interact (no interaction offset joins
    them), for odd and even edges;
  - the 216 classes' work units visit every site exactly once, every index stays
    inside the lattice and the per-colour lists fit their 64-entry buffers;
  - the Philox counters (line*32 + q, class) of the units are distinct.
"""
import itertools

import numpy as np
import pytest

import oracle as O

FCC_LIST = 64


def ring_colour(v, n):
    if n % 4 == 0 or v < n - 2:
        return v & 3
    return 4 if v == n - 2 else 5


def offsets():
    out = []
    for a, b, c in itertools.product(range(-2, 3), repeat=3):
        if (a + b + c) % 2 == 0 and a * a + b * b + c * c in (2, 4, 6):
            out.append((a, b, c))
    return out


@pytest.mark.parametrize("n", list(range(6, 258, 2)))
def test_ring_colouring(n):
    col = [ring_colour(v, n) for v in range(n)]
    for v in range(n):
        for d in (1, 2):
            assert col[v] != col[(v + d) % n]


def emulate_units(L):
    L0, L1, L2 = L
    lists = {}
    for axis, Lx in (("y", L1), ("z", L2)):
        for c in range(6):
            lists[axis, c] = [v for v in range(2 * Lx) if ring_colour(v, 2 * Lx) == c]
    for c in range(6):
        for p in (0, 1):
            lists["x", c, p] = [v >> 1 for v in range(2 * L0)
                                if ring_colour(v, 2 * L0) == c and (v & 1) == p]
    for key, v in lists.items():
        assert len(v) <= FCC_LIST, key
    visits, counters = [], set()
    for col in range(216):
        cx, cy, cz = col % 6, (col // 6) % 6, col // 36
        ny, nz = len(lists["y", cy]), len(lists["z", cz])
        maxq = (max(len(lists["x", cx, 0]), len(lists["x", cx, 1])) + 3) >> 2
        for u in range(ny * nz * maxq):
            q, li = u % maxq, u // maxq
            y, z = lists["y", cy][li % ny], lists["z", cz][li // ny]
            p = (y + z) & 1
            xs = lists["x", cx, p]
            if 4 * q >= len(xs):
                continue
            line = y + 2 * L1 * z
            key = (line * 32 + q, col)
            assert key not in counters
            counters.add(key)
            for s in range(4 * q, min(4 * q + 4, len(xs))):
                h = xs[s]
                assert 0 <= h < L0
                visits.append((col, line * L0 + h))
    return visits


@pytest.mark.parametrize("L", [(3, 3, 3), (4, 4, 4), (5, 3, 4), (6, 7, 3), (16, 16, 16),
                               (17, 17, 17), (127, 3, 3)])
def test_units_cover_every_site_once(L):
    N = 4 * L[0] * L[1] * L[2]
    visits = emulate_units(L)
    sites = sorted(s for _, s in visits)
    assert sites == list(range(N))


@pytest.mark.parametrize("L", [(3, 3, 3), (4, 4, 4), (5, 3, 4), (6, 7, 3)])
def test_classes_do_not_interact(L):
    x, y, z = O.fcc_coords(L)
    n = [2 * v for v in L]
    cls = np.array([ring_colour(a, n[0]) + 6 * ring_colour(b, n[1]) + 36 * ring_colour(c, n[2])
                    for a, b, c in zip(x, y, z)])
    index = {}
    for l, (a, b, c) in enumerate(zip(x, y, z)):
        index[a, b, c] = l
    for l, (a, b, c) in enumerate(zip(x, y, z)):
        for d in offsets():
            m = index[(a + d[0]) % n[0], (b + d[1]) % n[1], (c + d[2]) % n[2]]
            assert m != l and cls[m] != cls[l]
