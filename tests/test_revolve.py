"""The binomial checkpoint schedule of the torque adjoint (host code of the
library; no GPU needed): replaying it must visit every step's adjoint once,
in reverse order, each at the state it needs, within the snapshot budget and
the binomial bound on recomputation."""
import math

import pytest

import lbmflow_amd as L


def _replay(steps, snaps):
    sched = L.revolve_schedule(steps, snaps)
    slots = {0: 0}
    cur = None
    advanced = 0
    order = []
    for act, arg in sched:
        if act == "restore":
            assert arg in slots, f"restore of an empty slot {arg}"
            cur = slots[arg]
        elif act == "advance":
            assert cur is not None and arg >= 1
            cur += arg
            advanced += arg
        elif act == "save":
            assert 1 <= arg <= snaps, f"slot {arg} outside the budget of {snaps}"
            slots[arg] = cur
        else:
            assert cur == arg, f"adjoint of step {arg} at state {cur}"
            order.append(arg)
    return order, advanced


@pytest.mark.parametrize("steps,snaps", [(1, 0), (7, 0), (10, 2), (60, 3), (200, 5), (4000, 64), (333, 1)])
def test_schedule_reverses_every_step_once(steps, snaps):
    order, advanced = _replay(steps, snaps)
    assert order == list(range(steps - 1, -1, -1))
    if snaps == 0:
        assert advanced == steps * (steps - 1) // 2
    else:
        r = 1
        while math.comb(snaps + r, snaps) < steps:
            r += 1
        assert advanced <= r * steps, (advanced, r)


def test_schedule_rejects_bad_requests():
    with pytest.raises(L.LbmError):
        L.revolve_schedule(0, 3)
