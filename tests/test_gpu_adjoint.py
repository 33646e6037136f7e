"""Adjoint of the time-averaged shaft torque (lbm_adjoint_torque): D3Q19
TRT stirred tank with rotating IBM and a per-cell viscosity field.
Gradients against central finite differences of the same GPU forward, and
dJ/dOmega against a finite difference of the FP64 oracle."""
import numpy as np
import pytest

import cases
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N, STEPS, WINDOW = 24, 40, 15
NU0 = 0.01
TIP = 0.05


def _nu_field(tc, seed=3):
    x, y, z = cases.coords((N, N, N))
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 3)
    smooth = (np.sin(2 * np.pi * x / N + ph[0]) * np.cos(2 * np.pi * y / N + ph[1])
              * np.sin(2 * np.pi * z / N + ph[2]))
    return NU0 * (1.0 + 0.1 * smooth)


def _solver(tc, omega, nu):
    import lbmflow_amd as L

    s = L.Solver(19, (N, N, N), nu=NU0, collision="trt", solid=tc["solid"])
    s.set_omega_field((1.0 / (3.0 * nu + 0.5)).astype(np.float32))
    s.set_rotating_body(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, omega))
    return s


def _J(tc, omega, nu):
    s = _solver(tc, omega, nu)
    out = s.adjoint_torque(STEPS, WINDOW, 4, gradients=False)
    s.close()
    return out["J"]


@pytest.fixture(scope="module")
def tank():
    tc = cases.tank_case(N, nu=NU0, tip=TIP, collision="trt", Q=19)
    nu = _nu_field(tc)
    s = _solver(tc, tc["omega"], nu)
    res = s.adjoint_torque(STEPS, WINDOW, 3)
    s.close()
    return tc, nu, res


def test_adjoint_is_independent_of_the_checkpoint_budget(tank):
    tc, nu, res = tank
    for snaps in (0, 1, 12):
        s = _solver(tc, tc["omega"], nu)
        r2 = s.adjoint_torque(STEPS, WINDOW, snaps)
        s.close()
        assert r2["J"] == res["J"]
        assert r2["dJ_domega"] == res["dJ_domega"]
        assert np.array_equal(r2["dJ_dnu"], res["dJ_dnu"])
    assert res["forward_steps"] > STEPS


def test_domega_matches_central_difference(tank):
    tc, nu, res = tank
    w = tc["omega"]
    eps = 1e-2 * w
    fd = (_J(tc, w + eps, nu) - _J(tc, w - eps, nu)) / (2 * eps)
    ad = res["dJ_domega"]
    print(f"dJ/dOmega adjoint {ad:.8e}  central FD {fd:.8e}")
    assert abs(ad - fd) <= 1e-3 * abs(fd)


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


def _direction(tc):
    x, y, z = cases.coords((N, N, N))
    return NU0 * np.cos(2 * np.pi * (x + 2 * y) / N) * np.sin(2 * np.pi * z / N) * (tc["solid"] == 0)


def test_dnu_matches_f64_oracle_difference(tank):
    """The directional derivative along a smooth field against the FP64
    oracle's central difference (eps 1e-3: truncation ~1e-7 relative)."""
    tc, nu, res = tank
    d = _direction(tc)
    eps = 1e-3
    fd64 = (_J64(tc, nu + eps * d, tc["omega"]) - _J64(tc, nu - eps * d, tc["omega"])) / (2 * eps)
    ad = float(np.dot(res["dJ_dnu"], d))
    print(f"dJ/dnu . d adjoint {ad:.10e}  f64 oracle FD {fd64:.10e}")
    assert abs(ad - fd64) <= 1e-3 * abs(fd64)
    assert np.all(res["dJ_dnu"][tc["solid"] != 0] == 0.0)


def test_dnu_matches_gpu_central_difference(tank):
    """Same direction by central difference of the fp32 GPU forward.  A 1 %
    viscosity change moves J by ~1.5e-4 relative, close to the fp32
    trajectory's rounding noise, so this check is held to 5e-3."""
    tc, nu, res = tank
    d = _direction(tc)
    eps = 1e-2
    fd = (_J(tc, tc["omega"], nu + eps * d) - _J(tc, tc["omega"], nu - eps * d)) / (2 * eps)
    ad = float(np.dot(res["dJ_dnu"], d))
    print(f"dJ/dnu . d adjoint {ad:.8e}  fp32 GPU FD {fd:.8e}")
    assert abs(ad - fd) <= 5e-3 * abs(fd)


def test_domega_matches_f64_oracle_difference(tank):
    tc, nu, res = tank
    w = tc["omega"]
    eps = 1e-2 * w
    fd64 = (_J64(tc, nu, w + eps) - _J64(tc, nu, w - eps)) / (2 * eps)
    ad = res["dJ_domega"]
    print(f"dJ/dOmega adjoint {ad:.8e}  f64 oracle FD {fd64:.8e}")
    assert abs(ad - fd64) <= 1e-3 * abs(fd64)


def test_autograd_function_carries_the_adjoint(tank):
    """TankTorque (the torch.autograd.Function over lbm_adjoint_torque): J
    and the gradients that backward hands to torch equal the adjoint's own,
    through the chain rule of a loss built on J."""
    import torch

    from lbmflow_amd.autograd import TankTorque

    tc, nu, res = tank
    omega = torch.tensor(tc["omega"], dtype=torch.float64, requires_grad=True)
    nu_t = torch.tensor(nu.reshape(N, N, N), dtype=torch.float64, requires_grad=True)

    def make(w):
        return _solver(tc, w, nu), STEPS, WINDOW, 3

    J = TankTorque.apply(omega, nu_t, make)
    assert J.item() == res["J"]
    loss = 2.0 * J + J * J
    loss.backward()
    scale = 2.0 + 2.0 * res["J"]
    assert omega.grad.item() == pytest.approx(scale * res["dJ_domega"], rel=1e-15)
    assert nu_t.grad.shape == nu_t.shape
    np.testing.assert_allclose(nu_t.grad.numpy().reshape(-1), scale * res["dJ_dnu"], rtol=1e-15, atol=0)
    # no gradient requested: the forward runs without the reverse sweep
    with torch.no_grad():
        J2 = TankTorque.apply(omega.detach(), nu_t.detach(), make)
    assert J2.item() == pytest.approx(res["J"], rel=1e-12)
