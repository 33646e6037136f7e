"""Device-resident runner steps (lbm_run_ibm_span): the eager span and the
HIP-graph replay must reproduce the per-call step sequence exactly."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _make(tc, storage, scalar):
    import lbmflow_amd as L

    n = tc["n"]
    s = L.Solver(tc["Q"], (n, n, n), nu=tc["nu"], collision=tc["collision"], storage=storage, solid=tc["solid"])
    s.set_rotating_body(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
    k = None
    if scalar:
        k = s.enable_scalar_distribution(tc["nu"], "d3q7")
        s.set_scalar_concentration(k, tc["blob"])
    return s, k


def _state(s, k):
    c = None if k is None else s.gather_scalar_concentration(k)
    return s.populations(), c


@pytest.mark.parametrize("collision,Q,storage,scalar", [("cm", 27, "f32", True), ("trt", 19, "f16", False)])
def test_ibm_span_and_graph_match_per_step_calls(collision, Q, storage, scalar):
    tc = cases.tank_case(28, nu=0.01, collision=collision, Q=Q, sigma=3.0)
    a, ka = _make(tc, storage, scalar)
    b, kb = _make(tc, storage, scalar)
    g, kg = _make(tc, storage, scalar)
    done = 0
    # odd chunk lengths move the parity, so the graph is recaptured; reads
    # between chunks are the eager work interleaved with the replays
    for chunk in (4, 5, 2, 6, 1, 4):
        ta = []
        for _ in range(chunk):
            a.clear_body_force_field()
            ta.append(a.apply_rotating_ibm()["torque"])
            a.step()
        tb, db = b.run_with_ibm(chunk, graph=False)
        tg, dg = g.run_with_ibm(chunk, graph=True)
        done += chunk
        ta = np.array(ta)
        assert np.array_equal(ta, tb), f"eager span torque differs at step {done}"
        assert np.array_equal(ta, tg), f"graph span torque differs at step {done}"
        fa, ca = _state(a, ka)
        fb, cb = _state(b, kb)
        fg, cg = _state(g, kg)
        assert np.array_equal(fa, fb) and np.array_equal(fa, fg), f"populations differ at step {done}"
        if scalar:
            assert np.array_equal(ca, cb) and np.array_equal(ca, cg), f"scalar differs at step {done}"
        assert db["iterations"] == dg["iterations"]
    assert np.isfinite(fa).all()
