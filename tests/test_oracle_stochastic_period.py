"""Pins of the oracle's stochastic sampling step (SamplingParams.hh:193-204):
counts to the next sample are drawn one uniform per count until a draw falls
below the rate, so
  * they are geometric: P(dn = k) = r (1 - r)^(k - 1), mean 1/r, variance (1-r)/r^2;
  * a rate of 1 always gives 1 (every draw on [0, 1) is below 1);
  * the engine advances by exactly one 64-bit word per count (a 53-bit double
    takes one output of a 64-bit engine), which the engine's next output shows.
"""
import math

import numpy as np
import pytest

import oracle as O


def test_rate_one_gives_one_count_per_draw():
    dn, nxt = O.stochastic_count_steps(5, 1.0, 100)
    assert (dn == 1).all()
    assert nxt == O.mt64_outputs(5, 101)[-1]


@pytest.mark.parametrize("rate", [0.5, 0.1, 1.0 / 25.0])
def test_counts_are_geometric(rate):
    n = 20000
    dn, nxt = O.stochastic_count_steps(77, rate, n)
    assert dn.min() >= 1
    # one engine word per count
    assert nxt == O.mt64_outputs(77, int(dn.sum()) + 1)[-1]
    mean_sd = math.sqrt((1 - rate) / rate ** 2 / n)
    assert abs(dn.mean() - 1 / rate) < 5 * mean_sd
    # P(dn = 1) = rate, P(dn = 2) = rate (1 - rate)
    for k, p in ((1, rate), (2, rate * (1 - rate))):
        f = float(np.mean(dn == k))
        assert abs(f - p) < 5 * math.sqrt(p * (1 - p) / n), (k, f, p)


def test_a_planted_second_draw_per_count_would_show():
    # with two words per count the engine would sit sum(dn) words further on
    dn, nxt = O.stochastic_count_steps(3, 0.3, 50)
    assert nxt != O.mt64_outputs(3, 2 * int(dn.sum()) + 1)[-1]
