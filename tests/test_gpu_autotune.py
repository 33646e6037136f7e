"""The collide-stream autotuner (lbm_autotune_tiles): it times candidate
tile visit orders on a scratch copy, so the state it leaves is the state it
found, and the order it picks changes no result: a tuned solver and an
untuned one stay bit-identical over further steps of either parity."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_autotune_keeps_state_and_results(storage):
    import lbmflow_amd as L

    n = 40
    tc = cases.tank_case(n, nu=0.01, tip=0.05, collision="cm", Q=27)

    def build():
        s = L.Solver(27, (n, n, n), nu=0.01, collision="cm", storage=storage, solid=tc["solid"])
        s.set_rotating_body(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
        return s

    a, b = build(), build()
    for s in (a, b):
        s.run_with_ibm(7, graph=False)  # odd parity
    before = a.populations()
    t0 = a.time
    res = a.autotune_tiles(6)
    assert res["zb_even"] in (1, 2, 4, 8, 16, 32) and res["zb_odd"] in (1, 4, 8, 16, 32, 64)
    assert res["ms_even"] > 0 and res["ms_odd"] > 0
    assert a.time == t0
    assert np.array_equal(a.populations().view(np.uint32), before.view(np.uint32))
    for k, graph in [(3, False), (4, True)]:
        ta, _ = a.run_with_ibm(k, graph=graph)
        tb, _ = b.run_with_ibm(k, graph=graph)
        assert np.array_equal(ta, tb)
        assert np.array_equal(a.populations().view(np.uint32), b.populations().view(np.uint32))
    print(f"{storage}: {res}")


def test_autotune_rejects_bad_arguments():
    import lbmflow_amd as L

    n = 16
    solid, wall = cases.cavity_masks(n)
    s = L.Solver(19, (n, n, n), nu=0.02, solid=solid, wall_u=wall)
    with pytest.raises(L.LbmError):
        s.autotune_tiles(1)


@pytest.mark.parametrize("Q,collision,storage", [(27, "cm", "f32"), (19, "trt", "f32"), (27, "cm", "f16")])
def test_linear_tiles_match_row_tiles(Q, collision, storage):
    """Rows packed into 64-cell blocks of the linear index (no wrapping axis,
    row length not a multiple of 64) give the populations of the row tiles
    bit for bit, over odd and even spans, with the odd-step dead-slot fill
    (D3Q27 fp32) and a moving lid."""
    import os

    import lbmflow_amd as L

    n = 45
    tc = cases.tank_case(n, nu=0.01, tip=0.05, collision=collision, Q=Q)
    solid, wall = cases.cavity_masks(40, u_lid=0.06)

    def build(row):
        old = os.environ.get("LBM_ROW_TILES")
        os.environ["LBM_ROW_TILES"] = "1" if row else "0"
        try:
            s = L.Solver(Q, (n, n, n), nu=0.01, collision=collision, storage=storage, solid=tc["solid"])
            s.set_rotating_body(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
            c = L.Solver(Q, (40, 40, 40), nu=0.02, collision=collision, storage=storage, solid=solid, wall_u=wall)
        finally:
            if old is None:
                os.environ.pop("LBM_ROW_TILES", None)
            else:
                os.environ["LBM_ROW_TILES"] = old
        return s, c

    (a, ca), (b, cb) = build(False), build(True)
    for k, graph in [(3, False), (4, True), (5, False)]:
        ta, _ = a.run_with_ibm(k, graph=graph)
        tb, _ = b.run_with_ibm(k, graph=graph)
        assert np.array_equal(ta, tb)
        assert np.array_equal(a.populations().view(np.uint32), b.populations().view(np.uint32))
        ca.run(k)
        cb.run(k)
        assert np.array_equal(ca.populations().view(np.uint32), cb.populations().view(np.uint32))
