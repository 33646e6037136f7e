"""The torque adjoint over a longer horizon against the FP64 oracle: 64^3
D3Q19 TRT tank with the Rushton IBM and a smooth per-cell viscosity field,
1,000 steps, J = mean shaft torque over the last 250, sixteen binomial
snapshots.  dJ/dOmega and the directional dJ/dnu . d within 1e-3 relative
of the FP64 oracle's central differences."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

N, STEPS, WINDOW, SNAPS = 64, 1000, 250, 16
NU0, TIP = 0.01, 0.05


def _nu(tc):
    x, y, z = cases.coords((N, N, N))
    return NU0 * (1.0 + 0.1 * np.sin(2 * np.pi * x / N + 0.3) * np.cos(2 * np.pi * y / N + 1.1)
                  * np.sin(2 * np.pi * z / N + 2.0))


def _J64(tc, nu, om):
    s = O.OracleSolver(19, (N, N, N), prec=8, collision=O.TRT, nu=NU0, solid=tc["solid"])
    s.set_omega_field(1.0 / (3.0 * nu + 0.5))
    tq = []
    for _ in range(STEPS):
        s.clear_force_field()
        d = s.apply_ibm(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, om))
        s.step()
        tq.append(-d["torque"][2])
    return float(np.mean(tq[-WINDOW:]))


def test_adjoint_1000_steps_against_f64_oracle_differences():
    import lbmflow_amd as L

    tc = cases.tank_case(N, nu=NU0, tip=TIP, collision="trt", Q=19)
    nu = _nu(tc)
    s = L.Solver(19, (N, N, N), nu=NU0, collision="trt", solid=tc["solid"])
    s.set_omega_field((1.0 / (3.0 * nu + 0.5)).astype(np.float32))
    s.set_rotating_body(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
    res = s.adjoint_torque(STEPS, WINDOW, SNAPS)
    s.close()
    w = tc["omega"]
    eps = 1e-2 * w
    fd_w = (_J64(tc, nu, w + eps) - _J64(tc, nu, w - eps)) / (2 * eps)
    print(f"[adjoint64] dJ/dOmega adjoint {res['dJ_domega']:.8e}  f64 FD {fd_w:.8e}", flush=True)
    assert abs(res["dJ_domega"] - fd_w) <= 1e-3 * abs(fd_w)
    x, y, z = cases.coords((N, N, N))
    d = NU0 * np.cos(2 * np.pi * (x + 2 * y) / N) * np.sin(2 * np.pi * z / N) * (tc["solid"] == 0)
    e = 1e-3
    fd_n = (_J64(tc, nu + e * d, w) - _J64(tc, nu - e * d, w)) / (2 * e)
    ad = float(np.dot(res["dJ_dnu"], d))
    print(f"[adjoint64] dJ/dnu . d adjoint {ad:.10e}  f64 FD {fd_n:.10e}", flush=True)
    assert abs(ad - fd_n) <= 1e-3 * abs(fd_n)
