"""The adjoint's checkpointing schedule (host code of the library, no GPU) against the closed
form of binomial checkpointing (Griewank, "Achieving logarithmic growth of temporal and spatial
complexity in reverse automatic differentiation", 1992; Griewank & Walther, revolve, 2000):
reversing l steps with s snapshots costs t = r l - C(s + r, r - 1) forward steps, where r is the
repetition number with C(s + r - 1, s) < l <= C(s + r, s). The library's reverse sweep visits
`states` states (each reverse op consumes one stored or recomputed state) with `slots` stored
states, slot 0 holding the initial state: l = states, s = slots."""
from __future__ import annotations

from math import comb

import pytest


def griewank(l: int, s: int) -> int:
    r = 0
    while comb(s + r, s) < l:
        r += 1
    return 0 if r == 0 else r * l - comb(s + r, r - 1)


@pytest.mark.parametrize("states", [2, 3, 5, 10, 37, 64, 101, 250])
@pytest.mark.parametrize("slots", [2, 3, 4, 7, 12])
def test_schedule_cost_is_the_binomial_optimum(lbm, states, slots):
    assert lbm.checkpoint_forward_steps(states, slots) == griewank(states, slots)


def test_degenerate_budgets(lbm):
    # one slot (the initial state only): recompute from the start for every state
    for n in (2, 9, 40):
        assert lbm.checkpoint_forward_steps(n, 1) == n * (n - 1) // 2
    # enough slots for every state: one forward pass
    assert lbm.checkpoint_forward_steps(2001, 2001) == 2000
    assert lbm.checkpoint_forward_steps(2001, 5000) == 2000


def test_differentiation_instance_budget(lbm):
    # T = 2000 steps -> 2001 states, 45 ~ sqrt(T) slots: r = 3
    assert lbm.checkpoint_forward_steps(2001, 45) == 3 * 2001 - comb(48, 2)
