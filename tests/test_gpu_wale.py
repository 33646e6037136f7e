"""GPU parity of the WALE omega_plus producer (les.rs:120-211) against the
oracle (f32 restatement) on identical inputs.

Bars: the velocity gradient is assembled with the reference's operations in
f32 on both sides, so nu_t may differ only through the three pow() calls
(a few ulp): nu_t within 1e-5 relative (plus 1e-12 absolute for the cells
where it vanishes), omega within 1e-6 relative, clipped-cell counts equal up
to cells whose raw nu_t lies within that tolerance of the bound.  A run with
the per-step closure follows the oracle's update-then-step loop with the
populations within 2e-6 absolute of each other after 12 steps.
"""
import numpy as np
import pytest

import lbmflow_amd as L
from oracle import oracle as O
from tests import cases
from tests import wale_cases as W

pytestmark = pytest.mark.gpu

NUT_RTOL, NUT_ATOL, OMEGA_RTOL = 1e-5, 1e-12, 1e-6


def _tau(nu):
    return 1.0 / O.omegas(O.TRT, nu)[0]


def _compare(orc, hip, what, tau_eff_max=None):
    om_o, nut_o, k_o, mx_o = orc.wale(tau_eff_max=tau_eff_max)
    diag, nut_h = hip.wale_update(tau_eff_max=tau_eff_max, want_nu_t=True)
    om_h = hip.omega_field()
    assert om_h is not None, what
    np.testing.assert_allclose(nut_h, nut_o, rtol=NUT_RTOL, atol=NUT_ATOL, err_msg=what)
    np.testing.assert_allclose(om_h, om_o, rtol=OMEGA_RTOL, atol=0, err_msg=what)
    assert diag["max_nu_t_before_clipping"] == pytest.approx(mx_o, rel=NUT_RTOL, abs=NUT_ATOL), what
    if tau_eff_max is not None:
        # (tau_eff_max - tau0) / 3 in f32 as the reference's T: tau0 ~ 0.5
        # leaves f32 spacing 6e-8 on the difference
        nu0 = diag["nu_t_max"]
        assert nu0 == pytest.approx((tau_eff_max - (0.5 + 3 * hip.nu)) / 3.0, abs=1e-7)
        raw = orc.wale()[1]  # the oracle's unclipped values
        near = int(np.sum(np.abs(raw - nu0) <= 2 * NUT_RTOL * nu0))
        assert abs(diag["clipped_cells"] - k_o) <= near, what
        assert diag["clipped_fraction"] == pytest.approx(diag["clipped_cells"] / hip.n, rel=1e-6)
    return nut_o, diag


@pytest.mark.parametrize("Q", [19, 27])
def test_linear_field_in_an_open_box(Q):
    """Non-periodic faces (first-order differences) and the strain of the
    populations; every cell has the gradient G, so nu_t is uniform."""
    rng = np.random.default_rng(5)
    G = 2e-3 * rng.standard_normal((3, 3))
    dims, nu = (12, 9, 10), 0.08
    f, _, _ = W.linear_state(Q, dims, G, _tau(nu))
    f = f.astype(np.float32).astype(np.float64)
    orc, hip = cases.pair(Q, dims, nu=nu)
    orc.set_f(f)
    orc.update_moments()
    hip.set_populations(f)
    nut, _ = _compare(orc, hip, f"linear D3Q{Q}")
    np.testing.assert_allclose(nut, W.wale_closed_form(G), rtol=2e-3)


def test_moving_wall_channel():
    """One-sided wall differences through a moving half-way wall."""
    dims, nu = (8, 14, 6), 0.1
    f, solid, wall_u, _ = W.channel_state(19, dims, 2e-3, 3e-5, _tau(nu), True)
    # add a little cross-flow so every gradient entry is populated
    rng = np.random.default_rng(2)
    f = f + 1e-6 * rng.standard_normal(f.shape) * (solid == 0)
    f = f.astype(np.float32).astype(np.float64)
    orc, hip = cases.pair(19, dims, nu=nu, periodic=(1, 0, 1), solid=solid, wall_u=wall_u)
    orc.set_f(f)
    orc.update_moments()
    hip.set_populations(f)
    _compare(orc, hip, "channel")


@pytest.mark.parametrize("Q,steps", [(19, 7), (27, 8)])
def test_developed_cavity_both_parities(Q, steps):
    """A lid-driven cavity after an odd and an even number of steps (the two
    population layouts), with and without the tau_eff clip."""
    n = 18
    solid, wall = cases.cavity_masks(n, u_lid=0.08)
    orc, hip = cases.pair(Q, (n, n, n), nu=0.01, solid=solid, wall_u=wall)
    orc.step(steps)
    hip.run(steps)
    nut, _ = _compare(orc, hip, f"cavity D3Q{Q} {steps}")
    pos = nut[nut > 0]
    assert pos.size > 100
    tau_max = 0.5 + 3 * 0.01 + 3 * float(np.median(pos))
    _compare(orc, hip, f"cavity D3Q{Q} {steps} clipped", tau_eff_max=tau_max)


def test_force_field_enters_the_stress():
    """Uniform force plus a per-cell field (strain_rate_at's Guo term)."""
    dims = (16, 12, 10)
    rho, u = cases.tgv_field(dims, u0=0.03)
    orc, hip = cases.pair(19, dims, nu=0.02, periodic=(1, 1, 1), force=(1e-5, -2e-6, 3e-6))
    orc.init_with(rho, u)
    hip.init_with(rho, u)
    x, y, z = cases.coords(dims)
    ff = np.zeros((x.size, 3), np.float32)
    ff[:, 1] = 4e-5 * np.sin(0.3 * x)
    ff[:, 2] = -2e-5 * np.cos(0.5 * y)
    orc.set_force_field(ff.astype(np.float64).reshape(-1))
    hip.set_body_force_field(ff)
    orc.step(5)
    hip.run(5)
    _compare(orc, hip, "forced TGV")


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 0)])
def test_tall_column_spans_several_z_chunks(periodic, monkeypatch):
    """A domain taller than the gradient pass's z march (the default length
    for the size, then 64, 7 and 1 cells through LBM_WALE_ZLEN) with x and
    y extents that fill no tile: the march seams, the wrapped z neighbours
    of a periodic axis and the one-sided ends of a closed one, with an
    obstacle column crossing the seams.  Every march length gives the same
    field bit for bit."""
    dims = (70, 5, 150)
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[60:70, 2, 30:33] = 1
    solid[125:130, 1:3, 10] = 1
    if not periodic[0]:
        solid[:, :, 0] = solid[:, :, nx - 1] = 1
    if not periodic[2]:
        solid[0, :, :] = solid[nz - 1, :, :] = 1
    solid = solid.reshape(-1)
    rho, u = cases.tgv_field(dims, u0=0.03)
    orc, hip = cases.pair(19, dims, nu=0.01, periodic=periodic, solid=solid)
    orc.init_with(rho, u)
    hip.init_with(rho, u)
    orc.step(3)
    hip.run(3)
    nut, _ = _compare(orc, hip, f"tall column {periodic}")
    assert np.sum(nut > 0) > 1000
    om = hip.omega_field()
    for zlen in ("64", "7", "1"):
        monkeypatch.setenv("LBM_WALE_ZLEN", zlen)
        hip.wale_update()
        assert np.array_equal(hip.omega_field(), om), f"march length {zlen}"


def test_per_step_closure_follows_the_update_loop():
    """set_wale: every step recomputes omega from the pre-collision state,
    as WaleLes::update before each Solver::step does."""
    n = 16
    solid, wall = cases.cavity_masks(n, u_lid=0.1)
    orc, hip = cases.pair(19, (n, n, n), nu=0.005, solid=solid, wall_u=wall)
    orc.step(4)
    hip.run(4)
    hip.set_wale(True)
    for _ in range(12):
        om, _, _, _ = orc.wale()
        orc.set_omega_field(om)
        orc.step(1)
    hip.run(12)
    fl = solid == 0
    a = hip.populations()[:, fl].astype(np.float64)
    b = orc.get_f()[:, fl]
    assert np.abs(a - b).max() < 2e-6
    om_h = hip.omega_field()
    om_o, _, _, _ = orc.wale()
    # the installed field is the one of the last step's pre-collision state,
    # so it lags the current state by a step; both must be eddy-viscous
    assert np.all(om_h <= 1.0 / (3 * hip.nu + 0.5) + 1e-6)
    assert np.sum(om_h < 0.999 / (3 * hip.nu + 0.5)) > 50
    hip.set_wale(False)
    hip.set_omega_field(None)
    assert hip.omega_field() is None


def test_per_step_closure_in_a_captured_ibm_span():
    """With a rotating body the closure runs after the IBM spread of each
    step (the force field of the stress term is the new one); the captured
    span matches the uncaptured one bit for bit and the oracle's loop
    (clear field, IBM, WALE update, step) within the pow() round-off."""
    tc = cases.tank_case(32, collision="trt", Q=19, nu=0.004)
    n, steps = tc["n"], 6
    runs = []
    for graph in (False, True):
        s = L.Solver(19, (n, n, n), nu=tc["nu"], collision="trt", solid=tc["solid"])
        s.set_rotating_body(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
        s.set_wale(True)
        s.run_with_ibm(steps, graph=graph)
        runs.append(s.populations())
        s.close()
    assert np.array_equal(runs[0], runs[1])
    o = O.OracleSolver(19, (n, n, n), prec=4, collision=O.TRT, nu=tc["nu"], solid=tc["solid"])
    for _ in range(steps):
        o.clear_force_field()
        o.apply_ibm(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
        om, _, _, _ = o.wale()
        o.set_omega_field(om)
        o.step(1)
    fl = tc["solid"] == 0
    assert np.abs(runs[1][:, fl].astype(np.float64) - o.get_f()[:, fl]).max() < 2e-6


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_16bit_storage_field_is_bounded(storage):
    n = 18
    solid, wall = cases.cavity_masks(n, u_lid=0.08)
    hip = L.Solver(19, (n, n, n), nu=0.01, storage=storage, solid=solid, wall_u=wall)
    ref = L.Solver(19, (n, n, n), nu=0.01, storage="f32", solid=solid, wall_u=wall)
    hip.run(8)
    ref.run(8)
    _, nut = hip.wale_update(want_nu_t=True)
    _, nut_ref = ref.wale_update(want_nu_t=True)
    om = hip.omega_field()
    assert np.all(np.isfinite(om)) and np.all(om > 0) and np.all(om <= 1.0 / (3 * 0.01 + 0.5) * (1 + 1e-6))
    big = nut_ref > 0.1 * nut_ref.max()
    assert big.sum() > 20
    rel = np.abs(nut[big] - nut_ref[big]) / nut_ref[big]
    assert np.median(rel) < 0.1
