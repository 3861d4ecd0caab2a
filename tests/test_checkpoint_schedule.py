"""CPU checks of the checkpoint schedules of astaroth_amd/adjoint.py, executed on a symbolic
machine whose 'state' is the step index: every reverse step must see its own input state, steps
must be reversed from last to first, and revolve must not use more slots than granted."""
import pytest

from astaroth_amd import adjoint as AD


def execute(acts, steps, initial_slots):
    store = dict(initial_slots)
    cur = None
    reversed_steps = []
    forward = 0
    live = set()
    max_live = 0
    for a in acts:
        if a[0] == "restore":
            assert a[2] in store, a
            assert store[a[2]] == a[1], a
            cur = store[a[2]]
        elif a[0] == "advance":
            assert cur == a[1], a
            forward += a[2] - a[1]
            cur = a[2]
        elif a[0] == "store":
            assert cur == a[1], a
            store[a[2]] = cur
            if not (isinstance(a[2], tuple) and a[2][0] == "ckpt"):
                live.add(a[2])
                max_live = max(max_live, len(live))
        elif a[0] == "free":
            store.pop(a[1])
            live.discard(a[1])
        elif a[0] == "reverse":
            assert cur == a[1], a
            reversed_steps.append(a[1])
            cur = None
    assert reversed_steps == list(range(steps - 1, -1, -1))
    return forward, max_live


@pytest.mark.parametrize("steps,snaps", [(1, 0), (2, 1), (10, 0), (10, 2), (50, 3), (100, 5),
                                         (37, 4)])
def test_revolve_schedule(steps, snaps):
    acts = AD.revolve_actions(steps, snaps)
    fwd, live = execute(acts, steps, {"init": 0})
    assert live <= snaps
    cost, _ = AD._revolve_table(steps, snaps)
    assert fwd == cost[steps][snaps]
    # more slots never costs more recomputation
    if snaps > 0:
        assert cost[steps][snaps] <= cost[steps][snaps - 1]


def test_revolve_known_costs():
    cost, _ = AD._revolve_table(10, 3)
    assert cost[10][0] == 45          # recompute from the start for every step
    assert cost[2][1] == 1
    # with enough slots every step is advanced exactly once before it is reversed: l - 1
    big, _ = AD._revolve_table(16, 15)
    assert big[16][15] == 15


@pytest.mark.parametrize("steps,every", [(50, 10), (7, 3), (5, 5)])
def test_uniform_schedule(steps, every):
    acts = AD.uniform_actions(steps, every)
    ck = {("ckpt", s): s for s in range(0, steps, every)}
    fwd, _ = execute(acts, steps, ck)
    segs = [min(s + every, steps) - s for s in range(0, steps, every)]
    assert fwd == sum(l - 1 for l in segs)
