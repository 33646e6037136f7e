"""Batched domains (lbm_batch_*): stepping several tanks of different sizes,
speeds and viscosities together must give each member exactly what it
gets when run alone."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

SPECS = [(24, 0.04, 0.01), (30, 0.02, 0.005), (20, 0.06, 0.02)]  # (n, tip speed, nu)


def _member(n, tip, nu, Q, coll):
    import lbmflow_amd as L

    tc = cases.tank_case(n, nu=nu, tip=tip, collision=coll, Q=Q)
    s = L.Solver(Q, (n, n, n), nu=nu, collision=coll, solid=tc["solid"])
    s.set_rotating_body(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
    return s


@pytest.mark.parametrize("Q,coll", [(19, "trt"), (27, "cm")])
def test_batch_matches_members_run_alone(Q, coll):
    import lbmflow_amd as L

    alone = [_member(*sp, Q, coll) for sp in SPECS]
    members = [_member(*sp, Q, coll) for sp in SPECS]
    batch = L.Batch(members)
    info = batch.info()
    assert info["members"] == len(SPECS) and info["tiles"] > 0 and info["forced_cells"] > 0
    done = 0
    for chunk, graph in [(3, False), (4, True), (5, True), (2, False), (6, True)]:
        ta = np.stack([s.run_with_ibm(chunk, graph=False)[0] for s in alone])
        tb = batch.run(chunk, graph=graph)
        done += chunk
        assert np.array_equal(ta, tb), f"torque differs after {done} steps (graph={graph})"
        for k, (a, b) in enumerate(zip(alone, members)):
            assert np.array_equal(a.populations(), b.populations()), f"member {k} differs after {done} steps"
    # the autotuner works on copies: the members' state does not move
    before = [m.populations() for m in members]
    grid, ms = batch.autotune(4)
    assert grid > 0 and ms > 0.0
    for b0, m in zip(before, members):
        assert np.array_equal(b0, m.populations())
    ta = np.stack([s.run_with_ibm(4, graph=False)[0] for s in alone])
    assert np.array_equal(ta, batch.run(4, graph=True))
    for a, b in zip(alone, members):
        assert np.array_equal(a.populations(), b.populations())
    # a member changed behind the batch's back is refused
    members[1].set_masks(members[1].solid())
    with pytest.raises(L.LbmError):
        batch.run(1)
    batch.close()
