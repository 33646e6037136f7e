"""The plain-method configuration: 3D lid-driven cavity, D3Q19 TRT (magic
3/16), 64^3, fp32, Re 100, lid 0.05 lu, resting rims with the lid on the
ZPos rim (build_wall_rims, solver.rs:1357-1406; the reference's cavity test
t15_5_cavity3d.rs:148-181), quiescent start, 1000 steps.  The populations
are the oracle's bit for bit and so are the centreline profiles
(t15_5_cavity3d.rs:183-243); the profiles show the lid-driven vortex."""
import numpy as np
import pytest

import cases
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_cavity_64_re100_1000_steps_matches_oracle():
    import lbmflow_amd as L
    from lbmflow_amd import cavity

    n, u_lid, re, steps = 64, 0.05, 100.0, 1000
    nu = u_lid * (n - 2) / re
    solid, wall = cases.cavity_masks(n, u_lid=u_lid)
    wall32 = wall.astype(np.float32).astype(np.float64)
    print(f"[cavity] {n}^3 Re {re} {steps} steps", flush=True)
    hip = L.Solver(19, (n,) * 3, nu=nu, collision="trt", solid=solid, wall_u=wall32)
    hip.run(steps)
    orc = O.OracleSolver(19, (n,) * 3, prec=4, collision=O.TRT, nu=nu, solid=solid, wall_u=wall32)
    orc.step(steps)
    fl = solid == 0
    a = hip.populations()[:, fl].astype(np.float64)
    b = orc.get_f()[:, fl]
    assert np.array_equal(a, b), f"max |d| {np.abs(a - b).max()}"
    rho_h, ux_h, uy_h, uz_h = hip.moments()
    rho_o, ux_o, uy_o, uz_o = orc.moments()
    zh, uh, xh, wh = cavity.centrelines(ux_h, uz_h, n, u_lid)
    zo, uo, xo, wo = cavity.centrelines(ux_o, uz_o, n, u_lid)
    assert np.array_equal(uh, uo) and np.array_equal(wh, wo)
    # the primary vortex: forward flow under the lid, return flow below the
    # centre, downflow near x = 1 and upflow near x = 0 on the mid plane
    assert uh[-2] > 0.3 and uh[1:8].min() < -0.05
    assert wh[1:5].max() > 0.02 and wh[-6:-1].min() < -0.02
    # mass is conserved to round-off in a closed box
    assert abs(hip.mass_deviation()) <= 1e-6 * fl.sum()
