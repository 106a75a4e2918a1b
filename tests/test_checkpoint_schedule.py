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


# ---- pipelined multi-level schedule -------------------------------------------------------------
# Every launch of its reverse sweep holds one reverse step (plus one recomputation step per
# level), so the whole gradient is T forward launches then T + 1 reverse launches when no launch
# stalls. The library replays every schedule it builds on buffer contents (each step must read
# the state it expects, reverse steps must run T, ..., 0) and refuses one that fails.

@pytest.mark.parametrize("states", [2, 3, 10, 81, 500, 2001])
@pytest.mark.parametrize("slots", [6, 12, 20, 45, 91])
def test_pipelined_within_budget_or_binomial(lbm, states, slots):
    s = lbm.checkpoint_schedule(states, slots, "auto")
    assert s["slots"] <= max(slots, 1)
    if s["kind"] == "pipelined":
        T = states - 1
        assert 2 <= s["levels"] <= 4
        assert s["launches"] >= 2 * T + 1
        # one recomputation step per level and reverse step at most, plus the first pass
        assert T <= s["forward_steps"] <= T + (s["levels"] - 1) * states
    else:
        assert s["forward_steps"] == griewank(states, min(slots, states))


def test_differentiation_instance_pipelined_at_sqrt_T(lbm):
    # T = 2000, at most 45 ~ sqrt(T) stored states: three levels, no stalled launch
    s = lbm.checkpoint_schedule(2001, 45, "pipelined")
    assert s["kind"] == "pipelined" and s["levels"] == 3 and s["slots"] <= 45
    assert s["launches"] == 2 * 2000 + 1
    b = lbm.checkpoint_schedule(2001, 45, "binomial")
    assert b["forward_steps"] == 3 * 2001 - comb(48, 2)
    assert b["launches"] > s["launches"] + 2000


def test_pipelined_levels_follow_the_budget(lbm):
    # fewer stored states need more levels (more recomputation chains per launch)
    two = lbm.checkpoint_schedule(2001, 91, "pipelined")
    three = lbm.checkpoint_schedule(2001, 45, "pipelined")
    four = lbm.checkpoint_schedule(2001, 30, "pipelined")
    assert (two["levels"], three["levels"], four["levels"]) == (2, 3, 4)
    assert two["forward_steps"] < three["forward_steps"] < four["forward_steps"]


def test_pipelined_rejects_budgets_too_small(lbm):
    with pytest.raises(ValueError, match="pipelined"):
        lbm.checkpoint_schedule(2001, 6, "pipelined")
    assert lbm.checkpoint_schedule(2001, 6, "auto")["kind"] == "binomial"


def test_schedule_builders_under_sanitizers():
    """tests/cpp/test_schedule_sanitized.cpp: the library's schedule builders (host C++) built
    with AddressSanitizer and UndefinedBehaviorSanitizer, swept over thousands of (states, slots)
    pairs and level spacings; every schedule replayed on buffer contents, planted mistakes
    rejected, invalid arguments refused. Built by tests/cpp/Makefile (__graft_entry__.build())."""
    import os
    import pathlib
    import subprocess

    exe = pathlib.Path(__file__).resolve().parent / "cpp" / "bin" / "test_schedule_sanitized"
    if not exe.exists():
        pytest.skip("tests/cpp not built (make -C tests/cpp)")
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert " 0 failures" in r.stdout
