"""The binomial checkpoint schedule of the adjoint (host only, no device).

The schedule is executed symbolically: every action must find the state it
needs (a STORE / RECORD / ADVANCE at the step the working state holds, a
RESTORE from a slot holding that step), at most `snapshots` states are held,
each step is recorded right before its transpose, and the transposes run
n-1, ..., 0 exactly once.  The recomputation count must stay within the
fewest possible (Griewank and Walther's closed form r n - C(s + r, s + 1)
with r the least repetition number for which C(s + r, s) >= n).
"""
import math

import pytest

from eulerianfsi_amd import adjoint

ADVANCE, STORE, RESTORE, RECORD, ADJOINT, FREE = range(6)


def simulate(n, s):
    out = adjoint.plan(n, s, with_actions=True)
    cur, slots, expected, recorded, advances, records = 0, {}, n - 1, None, 0, 0
    forward_done = False
    for kind, step, slot, count in out["actions"]:
        if kind == STORE:
            assert cur == step and slot not in slots
            slots[slot] = step
            assert len(slots) <= s
        elif kind == RESTORE:
            assert slots[slot] == step
            cur = step
        elif kind == ADVANCE:
            assert cur == step and count >= 1
            cur += count
            advances += count
        elif kind == RECORD:
            assert cur == step
            recorded, cur = step, step + 1
            records += 1
            if not forward_done:
                assert step == n - 1  # the forward part ends with the last step, taped
                forward_done = True
        elif kind == ADJOINT:
            assert recorded == step == expected
            expected -= 1
            recorded = None
        elif kind == FREE:
            assert slots.pop(slot) == step
    assert expected == -1 and records == n
    assert advances == out["advances"]
    return out


def fewest(n, s):
    if n <= 1:
        return 0
    if s == 1:
        return n * (n - 1) // 2
    r = 0
    while math.comb(s + r, s) < n:
        r += 1
    return r * n - (math.comb(s + r, s + 1) if r >= 1 else 0)


@pytest.mark.parametrize("n,s", [(1, 1), (2, 1), (7, 1), (10, 2), (20, 3), (100, 5), (2000, 45), (64, 64),
                                 (2000, 2000), (37, 4)])
def test_schedule_reverses_every_step_once(n, s):
    out = simulate(n, s)
    r = out["repetitions"]
    assert out["advances"] <= r * n
    assert out["advances"] == fewest(n, s)
    if s == 1:
        assert out["advances"] == n * (n - 1) // 2
    if s >= n:
        assert out["advances"] == n - 1 and r == 1


def test_config_schedule_is_binomial():
    # 2000 steps with 45 snapshots: repetition number 3 (C(45 + 2, 45) = 1081 < 2000 <= C(45 + 3, 45)),
    # 3 * 2000 - C(48, 46) = 4872 recomputed steps
    out = simulate(2000, 45)
    assert out["repetitions"] == 3
    assert math.comb(47, 45) < 2000 <= math.comb(48, 45)
    assert out["most_held"] <= 45
    assert out["advances"] == 4872
