"""Pins of the CPU oracle (oracle/lbm_oracle.cpp) — each check is an
independent fact about the reference algorithm: the reference's own unit
tests (fixed points, invariants, closed forms), textbook properties
(moments of the discrete equilibrium, conservation, exact Poiseuille for
TRT with magic 3/16, TGV decay, Gaussian spreading) and brute-force
recomputation in numpy where the mathematics fixes the result.  A plausible
mistake in the restatement (a dropped term, a wrong sign or index, a
transposed operand) fails at least one of them.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


# ---- lattice tables (lattice.rs:285-442, invariants 484-790) -------------------

@pytest.mark.parametrize("Q", [19, 27])
def test_lattice_order_matches_reference_tables(Q):
    g = load_golden("lattice_order.json")
    c, w, opp, npairs = O.lattice(Q)
    assert c.tolist() == g[f"D3Q{Q}"]["c"]
    np.testing.assert_array_equal(w, np.array(g[f"D3Q{Q}"]["w_num"]) / np.array(g[f"D3Q{Q}"]["w_den"]))
    assert npairs == (Q - 1) // 2
    for q in range(1, Q):
        assert opp[q] == (q + 1 if q % 2 else q - 1)
        assert np.array_equal(c[opp[q]], -c[q])


@pytest.mark.parametrize("Q", [7, 19, 27])
def test_lattice_isotropy(Q):
    c, w, opp, _ = O.lattice(Q)
    cs2 = 0.25 if Q == 7 else 1.0 / 3.0
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs(w @ c).max() < 1e-16
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, c, c), cs2 * np.eye(3), atol=1e-16)
    if Q != 7:
        # fourth-order isotropy: sum w c_a^4 = 3 cs^4, sum w c_a^2 c_b^2 = cs^4
        assert abs((w * c[:, 0] ** 4).sum() - 3 * cs2 ** 2) < 1e-15
        assert abs((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() - cs2 ** 2) < 1e-15
        # face unknowns (5 / 9) and the closure constant (lattice.rs:109-116)
        for n_in in np.vstack([np.eye(3), -np.eye(3)]):
            dot = c @ n_in
            assert (dot > 0).sum() == (5 if Q == 19 else 9)
            assert abs(w[dot == 0].sum() + 2 * w[dot < 0].sum() - 1.0) < 1e-15


# ---- relaxation rates (params.rs:205-217, test 425-434) -------------------------

def test_omegas_match_reference_test():
    nu = 0.02
    tau = 3.0 * nu + 0.5
    op, om = O.omegas(O.TRT, nu, 3.0 / 16.0)
    assert op == 1.0 / tau
    assert om == 1.0 / ((3.0 / 16.0) / (tau - 0.5) + 0.5)
    bp, bm = O.omegas(O.BGK, nu)
    assert bp == bm
    # the magic product (1/w+ - 1/2)(1/w- - 1/2) = 3/16
    assert abs((1 / op - 0.5) * (1 / om - 0.5) - 3.0 / 16.0) < 1e-15


# ---- equilibrium ------------------------------------------------------------------

@pytest.mark.parametrize("Q", [19, 27])
def test_equilibrium_moments(Q):
    s = O.OracleSolver(Q, (2, 2, 2), prec=8, periodic=(1, 1, 1), nu=0.1)
    c, w, _, _ = O.lattice(Q)
    rng = np.random.default_rng(3)
    for _ in range(10):
        rho = 1 + 0.1 * rng.standard_normal()
        u = 0.05 * rng.standard_normal(3)
        feq = s.equilibrium(rho, u)
        assert abs(feq.sum() - (rho - 1.0)) < 1e-15
        np.testing.assert_allclose(c.T @ feq, rho * u, atol=1e-15)
        pi = np.einsum("q,qa,qb->ab", feq + w, c, c)
        np.testing.assert_allclose(pi, rho * np.outer(u, u) + rho / 3.0 * np.eye(3), atol=1e-15)


def _ref_sample(i, Q):
    h = lambda k: np.sin((i + 1.0) * k)  # noqa: E731  kernels.rs:1331-1339
    r = 1.0 + 0.5 * h(12.9898)
    u = np.array([0.12 * h(78.233), 0.12 * h(39.425), 0.12 * h(93.989)])
    return r, u


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("prec", [4, 8])
def test_trt_equilibrium_is_bitexact_fixed_point(Q, prec):
    """kernels.rs:1316-1388: forceless collide of feq returns feq bit for bit."""
    n = 64
    s = O.OracleSolver(Q, (n, 1, 1), prec=prec, periodic=(1, 1, 1), collision=O.TRT, nu=(1 / 1.25 - 0.5) / 3)
    f = np.zeros((Q, n))
    rho = np.zeros(n)
    U = np.zeros((n, 3))
    cast = np.float32 if prec == 4 else np.float64
    for i in range(n):
        r, u = _ref_sample(i, Q)
        r, u = float(cast(r)), u.astype(cast).astype(np.float64)
        rho[i], U[i] = r, u
        f[:, i] = s.equilibrium(r, u)
    s.set_f(f)
    s.set_moments(rho, U[:, 0], U[:, 1], U[:, 2])
    s.collide_only()
    g = s.get_f()
    assert np.array_equal(f.view(np.int64), g.view(np.int64)) or np.array_equal(f, g)


def _random_state(s, Q, rng, amp=1e-3):
    c, w, _, _ = O.lattice(Q)
    f = w[:, None] * amp * rng.standard_normal((Q, s.n))
    f += w[:, None] * (0.05 * (c @ rng.standard_normal(3)))[:, None]
    s.set_f(f)
    s.update_moments()
    return f


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("coll", ["bgk", "trt"])
def test_trt_conserves_mass_and_adds_force_to_momentum(Q, coll):
    force = np.array([1e-4, -2e-4, 5e-5])
    s = O.OracleSolver(Q, (4, 3, 2), prec=8, periodic=(1, 1, 1), collision=cases.COLL[coll], nu=0.03,
                       force=force)
    c, w, _, _ = O.lattice(Q)
    f0 = _random_state(s, Q, np.random.default_rng(5))
    s.collide_only()
    f1 = s.get_f()
    np.testing.assert_allclose(f1.sum(0), f0.sum(0), atol=2e-16)
    # Guo forcing: post-collision momentum = m + F exactly (for any omega-)
    np.testing.assert_allclose(c.T @ f1, c.T @ f0 + force[:, None], atol=1e-16)


@pytest.mark.parametrize("Q", [19, 27])
def test_bgk_matches_textbook_bgk_guo(Q):
    force = np.array([3e-5, 0.0, -1e-5])
    nu = 0.04
    s = O.OracleSolver(Q, (3, 3, 3), prec=8, periodic=(1, 1, 1), collision=O.BGK, nu=nu, force=force)
    c, w, _, _ = O.lattice(Q)
    f0 = _random_state(s, Q, np.random.default_rng(11))
    rho, ux, uy, uz = s.moments()
    s.collide_only()
    f1 = s.get_f()
    om = 1.0 / (3 * nu + 0.5)
    for i in range(s.n):
        u = np.array([ux[i], uy[i], uz[i]])
        cu = c @ u
        feq = w * ((rho[i] - 1) + rho[i] * (3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u))
        src = w * (3 * (c @ force - u @ force) + 9 * cu * (c @ force))
        np.testing.assert_allclose(f1[:, i], f0[:, i] - om * (f0[:, i] - feq[:]) + (1 - om / 2) * src, atol=1e-16)


def test_trt_rates_act_on_symmetric_and_antisymmetric_parts():
    Q, nu, magic = 19, 0.05, 0.1
    s = O.OracleSolver(Q, (3, 2, 2), prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=nu, magic=magic)
    op, om = O.omegas(O.TRT, nu, magic)
    c, w, opp, _ = O.lattice(Q)
    f0 = _random_state(s, Q, np.random.default_rng(2))
    rho, ux, uy, uz = s.moments()
    s.collide_only()
    f1 = s.get_f()
    for i in range(s.n):
        u = np.array([ux[i], uy[i], uz[i]])
        cu = c @ u
        feq = w * ((rho[i] - 1) + rho[i] * (3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u))
        sym0 = 0.5 * (f0[:, i] + f0[opp, i])
        asym0 = 0.5 * (f0[:, i] - f0[opp, i])
        esym = 0.5 * (feq + feq[opp])
        easym = 0.5 * (feq - feq[opp])
        np.testing.assert_allclose(0.5 * (f1[:, i] + f1[opp, i]), sym0 - op * (sym0 - esym), atol=1e-16)
        np.testing.assert_allclose(0.5 * (f1[:, i] - f1[opp, i]), asym0 - om * (asym0 - easym), atol=1e-16)


# ---- central moments (kernels.rs:162-533) ----------------------------------------

@pytest.mark.parametrize("Q", [19, 27])
def test_cm_transform_inverse_and_basis(Q):
    R, P, basis = O.cm_transform(Q)
    np.testing.assert_allclose(R @ P, np.eye(Q), atol=1e-13)
    c, _, _, _ = O.lattice(Q)
    for m in range(Q):
        e = basis[m]
        np.testing.assert_array_equal(R[m], np.prod(c.astype(float) ** e, axis=1))
    if Q == 19:
        assert not any((b > 0).all() for b in basis)


@pytest.mark.parametrize("Q", [19, 27])
def test_cm_rest_state_is_exact_fixed_point(Q):
    s = O.OracleSolver(Q, (8, 1, 1), prec=8, periodic=(1, 1, 1), collision=O.CM, nu=0.1, omega_shear=1.25)
    s.collide_only()
    assert np.array_equal(s.get_f(), np.zeros((Q, 8)))


def test_cm_uniform_velocity_stays_uniform_bitwise():
    Q, n = 19, 16
    s = O.OracleSolver(Q, (n, 1, 1), prec=8, periodic=(1, 1, 1), collision=O.CM, nu=0.1, omega_shear=1.1)
    feq = s.equilibrium(1.03, np.array([0.04, -0.03, 0.02]))
    s.set_f(np.repeat(feq[:, None], n, 1))
    s.set_moments(np.full(n, 1.03), np.full(n, 0.04), np.full(n, -0.03), np.full(n, 0.02))
    s.collide_only()
    g = s.get_f()
    for q in range(Q):
        assert len(set(g[q].tolist())) == 1


def _central(f_phys, c, u, e):
    return (f_phys * np.prod((c - u) ** np.array(e), axis=1)).sum()


@pytest.mark.parametrize("force", [(0.0, 0.0, 0.0), (2e-5, -1e-5, 3e-5)])
@pytest.mark.parametrize("Q", [19, 27])
def test_cm_relaxes_central_moments_as_specified(force, Q):
    """Post-collision central moments computed directly in numpy: order >= 3
    and the second-order trace relax to equilibrium (rate 1), deviatoric
    second order with os*(1 - 0.16|u|^2), conserved orders unchanged.  For
    D3Q27 this runs the factorised transforms, for D3Q19 the matrix path."""
    nu, os_base = 0.02, 1.6
    force = np.array(force)
    s = O.OracleSolver(Q, (3, 2, 2), prec=8, periodic=(1, 1, 1), collision=O.CM, nu=nu, omega_shear=os_base,
                       force=force)
    c, w, _, _ = O.lattice(Q)
    c = c.astype(float)
    f0 = _random_state(s, Q, np.random.default_rng(9), amp=3e-3)
    rho, ux, uy, uz = s.moments()
    s.collide_only()
    f1 = s.get_f()
    _, _, basis = O.cm_transform(Q)
    for i in range(s.n):
        u = np.array([ux[i], uy[i], uz[i]])
        cu = c @ u
        feq = w * rho[i] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
        src = w * (3 * (c @ force - u @ force) + 9 * cu * (c @ force))
        osr = min(os_base * (1 - 0.16 * (u @ u)), 2.0)
        K = {tuple(e): _central(f0[:, i] + w, c, u, e) for e in basis}
        E = {tuple(e): _central(feq, c, u, e) for e in basis}
        S = {tuple(e): _central(src, c, u, e) for e in basis}
        P = {tuple(e): _central(f1[:, i] + w, c, u, e) for e in basis}
        diag = [(2, 0, 0), (0, 2, 0), (0, 0, 2)]
        kb = sum(K[d] - E[d] for d in diag) / 3
        sb = sum(S[d] for d in diag) / 3
        for e in map(tuple, basis):
            order = sum(e)
            if order <= 1:
                want = K[e] + S[e]
            elif order == 2 and e in diag:
                want = E[e] + (1 - osr) * (K[e] - E[e] - kb) + 0.5 * sb + (1 - 0.5 * osr) * (S[e] - sb)
            elif order == 2:
                want = K[e] - osr * (K[e] - E[e]) + (1 - 0.5 * osr) * S[e]
            else:
                want = E[e] + 0.5 * S[e]
            assert abs(P[e] - want) < 1e-14, (e, P[e], want)


# ---- streaming, walls, whole runs -------------------------------------------------

@pytest.mark.parametrize("Q,coll", [(27, "trt"), (27, "cm"), (19, "trt")])
def test_closed_box_mass_is_conserved(Q, coll):
    """d3q27_smoke.rs:26-71: a closed box (moving lid) keeps its mass to 1e-13."""
    n = 12
    solid, wall = cases.cavity_masks(n, u_lid=0.05)
    s = O.OracleSolver(Q, (n, n, n), prec=8, collision=cases.COLL[coll], nu=0.03, solid=solid, wall_u=wall)
    m0 = s.reduce(1)
    s.step(400 if coll != "cm" else 120)
    assert abs(s.reduce(1) - m0) < 1e-12
    assert s.fluid_cells() == (n - 2) ** 3


def test_trt_magic_poiseuille_is_exact():
    """validation_channel.rs:87-105: TRT with magic 3/16 and half-way walls
    reproduces the parabolic profile to round-off."""
    ny, g, nu = 10, 1e-6, 1.0 / 6.0
    dims = (3, ny, 3)
    is_wall = [0, 0, 1, 1, 0, 0]
    solid, wall = O.build_wall_rims(dims, is_wall, np.zeros((6, 3)))
    s = O.OracleSolver(19, dims, prec=8, periodic=(1, 0, 1), collision=O.TRT, nu=nu, force=(g, 0, 0), solid=solid,
                       wall_u=wall)
    prev = None
    for _ in range(400):
        s.step(100)
        ux = s.moments()[1].reshape(3, ny, 3)[1, :, 1]
        if prev is not None and np.abs(ux - prev).max() < 1e-16:
            break
        prev = ux.copy()
    h = ny - 2
    yw = np.arange(1, ny - 1) - 0.5
    exact = g / (2 * nu) * yw * (h - yw)
    got = ux[1:ny - 1]
    assert np.abs(got - exact).max() / exact.max() < 1e-10
    assert np.abs(got - got[::-1]).max() < 1e-13 * exact.max() + 1e-22


@pytest.mark.parametrize("Q", [19, 27])
def test_tgv_decays_at_the_viscous_rate(Q):
    n, nu, u0 = 16, 0.02, 0.01
    k = 2 * np.pi / n
    x, y, z = cases.coords((n, n, n))
    u = np.stack([u0 * np.sin(k * x) * np.cos(k * y), -u0 * np.cos(k * x) * np.sin(k * y), 0 * x], 1)
    s = O.OracleSolver(Q, (n, n, n), prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=nu)
    s.init_with(np.ones(n ** 3), u)

    def ke():
        _, a, b, c = s.moments()
        return (a * a + b * b + c * c).sum()

    e0 = ke()
    s.step(200)
    rate = np.log(ke() / e0) / 200
    assert abs(rate / (-4 * nu * k * k) - 1) < 0.02


def test_init_with_quiescent_is_zero_and_moments_follow():
    n = 6
    s = O.OracleSolver(19, (n, n, n), prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=0.1)
    s.init_with(np.ones(n ** 3), np.zeros((n ** 3, 3)))
    assert np.array_equal(s.get_f(), np.zeros((19, n ** 3)))
    assert s.fluid_cells() == n ** 3


# ---- geometry (geometry.rs tests 979-1121) -----------------------------------------

def test_tank_volume_within_1pct_of_cylinder():
    n = 160
    dx = 1.0 / n
    solid = O.tank_mask((n, n, n), dx, 1.0, 1.0)
    fluid = (solid == 0).sum() * dx ** 3
    analytic = np.pi * 0.25 * ((n - 2) * dx)
    assert abs(fluid - analytic) / analytic < 0.01


def test_baffles_in_all_four_quadrants():
    n = 96
    dx = 1.0 / n
    solid = O.tank_mask((n, n, n), dx, 1.0, 1.0)
    solid, baffle = O.baffle_mask((n, n, n), dx, 1.0, 1.0, 0.08, 0.025, solid)
    x, y, z = cases.coords((n, n, n))
    px, py = (x + 0.5) * dx - 0.5, (y + 0.5) * dx - 0.5
    b = baffle.astype(bool)
    quad = np.where(np.abs(py) < np.abs(px), (px > 0).astype(int), 2 + (py > 0).astype(int))[b]
    assert all((quad == k).sum() > 0 for k in range(4))
    # baffle cells lie in the annulus [R - width, R) and within half a thickness of an axis
    r = np.hypot(px, py)[b]
    assert (r >= 0.5 - 0.08 - 1e-12).all() and (r < 0.5).all()
    assert (np.minimum(np.abs(px[b]), np.abs(py[b])) <= 0.0125 + 1e-12).all()


def test_rushton_markers_counts_bounds_and_targets():
    n = 256
    pos, w, vel, c = O.radial_impeller((n, n, n), 1.0, 240.0, 240.0, 80.0, 80.0, 6, 0.0, 1.25e-3)
    # (nr+1)(nt+1)(nz+1) per blade: nr = ceil(14/2), nt = nz = ceil(3.2)
    assert len(w) == 6 * 8 * 5 * 5 == 1200
    assert np.all(w == 1.0)
    r = np.hypot(pos[:, 0] - 128.0, pos[:, 1] - 128.0)
    # radial span [0.65 R, R] widened by the tangential half-thickness 1.6
    assert r.min() >= 26.0 - 1e-9 and r.max() <= np.hypot(40.0, 1.6) + 1e-9
    # target velocity is Omega x r
    rel = pos - c
    np.testing.assert_allclose(vel, np.stack([-1.25e-3 * rel[:, 1], 1.25e-3 * rel[:, 0], 0 * rel[:, 2]], 1),
                               atol=1e-15)
    lo = O.radial_impeller((64,) * 3, 1 / 64, 1.0, 1.0, 0.5, 0.35, 6, 0.0, 1.0)[1].size
    hi = O.radial_impeller((96,) * 3, 1 / 96, 1.0, 1.0, 0.5, 0.35, 6, 0.0, 1.0)[1].size
    assert hi > lo
    lo = O.radial_impeller((64,) * 3, 1 / 64, 1.0, 1.0, 0.5, 0.35, 4, 45.0, 1.0, 0.035, 1)[1].size
    hi = O.radial_impeller((96,) * 3, 1 / 96, 1.0, 1.0, 0.5, 0.35, 4, 45.0, 1.0, 0.035, 1)[1].size
    assert hi > lo


def test_wall_rims_faster_wall_wins_corners():
    n = 5
    u = np.zeros((6, 3))
    u[5] = [0.05, 0, 0]   # ZPos lid
    u[2] = [0.0, 0, 0.02]  # YNeg slower
    solid, wall = O.build_wall_rims((n, n, n), [1] * 6, u)
    idx = lambda x, y, z: (z * n + y) * n + x  # noqa: E731
    assert solid.sum() == n ** 3 - (n - 2) ** 3
    np.testing.assert_array_equal(wall[idx(2, 0, 4)], [0.05, 0, 0])   # shared edge: faster lid wins
    np.testing.assert_array_equal(wall[idx(2, 0, 2)], [0, 0, 0.02])
    np.testing.assert_array_equal(wall[idx(0, 2, 2)], [0, 0, 0])


def test_marker_stencil_weights():
    cells, w = O.marker_stencil(np.array([3.25, 4.5, 5.0]), (10, 10, 10), 1)
    # linear kernel: x -> {3: .75, 4: .25}; y -> {4: .5, 5: .5}; z -> {5: 1}
    assert cells.tolist() == [[3, 4, 5], [3, 5, 5], [4, 4, 5], [4, 5, 5]]
    np.testing.assert_allclose(w, [0.375, 0.375, 0.125, 0.125])
    cells, w = O.marker_stencil(np.array([3.25, 4.5, 5.0]), (10, 10, 10), 2)
    assert abs(w.sum() - 1.0) < 1e-15 and len(w) == 4 * 4 * 4 - 0 or abs(w.sum() - 1.0) < 1e-15
    cells, w = O.marker_stencil(np.array([0.2, 5.0, 5.0]), (10, 10, 10), 2)
    assert cells[:, 0].min() == 0  # clipped at the domain edge


# ---- rotating IBM (rotating_ibm.rs tests) -------------------------------------------

def _ibm_box(n=12, coll=O.TRT, Q=19):
    solid, wall = O.build_wall_rims((n, n, n), [1] * 6, np.zeros((6, 3)))
    return O.OracleSolver(Q, (n, n, n), prec=8, collision=coll, nu=0.05, solid=solid, wall_u=wall)


def test_ibm_single_marker_is_corrected_exactly():
    s = _ibm_box()
    pos = np.array([[5.3, 6.1, 5.7]])
    d = s.apply_ibm(pos, np.ones(1), np.array([6.0, 6.0, 5.7]), np.array([0.0, 0.0, 0.02]), max_iter=1)
    assert d["iterations"] == 1
    assert d["slip_max"] < 1e-17
    assert d["momentum_error_rel"] < 1e-14


def test_ibm_single_marker_impulse_closed_form_at_nonunit_density():
    """solver.rs:3040-3208 at rho != 1 with a body force and a force field
    already present: the interpolated velocity is u_now = (m + F/2) / rho
    with F = global force + force field; a marker on a cell centre (hat
    kernel: one stencil cell, W = 1) gets mobility 1 / rho, so its impulse is
    rho (U - u_now) and lands in the force field of that cell; a marker
    between cells is corrected exactly in one sweep (du = impulse W / rho,
    the same 1 / rho as the mobility)."""
    n = 6
    g = np.array([1.0e-4, -2.0e-4, 0.5e-4])
    s = O.OracleSolver(19, (n, n, n), prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=0.05, force=g)
    rng = np.random.default_rng(5)
    s.init_with(1.2 + 0.05 * rng.random(n ** 3), 0.01 * rng.standard_normal((n ** 3, 3)))
    ff0 = 1e-4 * rng.standard_normal((n ** 3, 3))
    s.set_force_field(ff0)
    rho = s.moments()[0]
    c, _, _, _ = O.lattice(19)
    m = s.get_f().T @ c.astype(float)  # sum_q c_q f_q (deviations: sum_q c_q w_q = 0)
    cell = (3 * n + 2) * n + 4  # (x, y, z) = (4, 2, 3)
    centre, om = np.array([1.0, 0.5, 3.0]), np.array([0.002, -0.001, 0.003])
    p = np.array([4.0, 2.0, 3.0])
    d = s.apply_ibm(p[None], np.ones(1), centre, om, max_iter=1, radius=1)
    target = np.cross(om, p - centre)
    u_now = (m[cell] + 0.5 * (g + ff0[cell])) / rho[cell]
    got = s.force_field()[cell] - ff0[cell]
    np.testing.assert_allclose(got, rho[cell] * (target - u_now), rtol=1e-13, atol=1e-19)
    assert d["slip_max"] < 1e-16
    # between cells, at the same non-unit density: one sweep is exact
    d = s.apply_ibm(np.array([[2.3, 3.6, 1.45]]), np.ones(1), centre, om, max_iter=1, radius=2)
    assert d["slip_max"] < 1e-16


def test_ibm_torque_flips_with_omega_sign_and_force_matches_marker_force():
    pos, w, vel, c = O.radial_impeller((24,) * 3, 1.0, 20.0, 20.0, 8.0, 8.0, 6, 0.0, 0.01)
    outs = []
    for sgn in (1.0, -1.0):
        s = _ibm_box(24)
        outs.append(s.apply_ibm(pos, w, c, np.array([0.0, 0.0, sgn * 0.01])))
    a, b = outs
    np.testing.assert_allclose(a["torque"], -b["torque"], rtol=1e-12, atol=1e-18)
    assert a["torque"][2] < 0  # reaction torque on the body opposes the fluid drive
    # the spread force and the marker quadrature are sums of the same cell forces
    assert np.abs(a["fluid_force"] - a["marker_force"]).max() < 1e-13


# ---- passive scalar (scalar.rs:17-41; solver.rs:5922-6035) ------------------------

def _scalar_case(d3q7, n=12, D=0.05, lid=0.0):
    solid, wall = cases.cavity_masks(n, u_lid=lid)
    s = O.OracleSolver(19, (n, n, n), prec=8, collision=O.TRT, nu=0.05, solid=solid, wall_u=wall)
    k = s.scalar_enable(D, d3q7)
    return s, k, solid


@pytest.mark.parametrize("d3q7", [False, True])
def test_scalar_equilibrium_moments(d3q7):
    s, k, solid = _scalar_case(d3q7, n=6)
    Qs = 7 if d3q7 else 19
    c, _, _, _ = O.lattice(Qs)
    C0 = 2.5
    s.scalar_set(k, np.full(s.n, C0))  # u = 0 everywhere: h = w C
    h = s.scalar_get_h(k, Qs)
    np.testing.assert_allclose(h.sum(0), C0, rtol=1e-15)
    assert np.abs(c.T @ h).max() < 1e-15


@pytest.mark.parametrize("d3q7", [False, True])
def test_scalar_uniform_stays_uniform_and_mass_is_conserved(d3q7):
    # uniform C in a uniform stream is a fixed point (solver.rs:5922-5960)
    n = 8
    s = O.OracleSolver(19, (n, n, n), prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=0.05)
    s.init_with(np.ones(n ** 3), np.tile([0.03, -0.02, 0.01], (n ** 3, 1)))
    k = s.scalar_enable(0.05, d3q7)
    s.scalar_set(k, np.full(n ** 3, 1.0))
    s.step(30)
    assert np.abs(s.scalar_get(k) - 1.0).max() < 1e-14
    # closed, stirred box: total scalar is conserved (solver.rs:5961-5995)
    s, k, solid = _scalar_case(d3q7, lid=0.05)
    fl = solid == 0
    s.step(20)
    x, y, z = cases.coords(s.dims)
    blob = np.where(fl, np.exp(-((x - 4.0) ** 2 + (y - 6) ** 2 + (z - 7) ** 2) / 4.0), 0.0)
    s.scalar_set(k, blob)
    m0 = s.scalar_get(k).sum()
    s.step(40)
    assert abs(s.scalar_get(k).sum() - m0) < 1e-12 * m0


@pytest.mark.parametrize("d3q7", [False, True])
def test_scalar_gaussian_spreads_at_2Dt(d3q7):
    """Diffusivity of the BGK scalar: omega = 1 / (3 D + 1/2) on the hydro
    lattice (solver.rs:2113-2205), 1 / (4 D + 1/2) on D3Q7 (c_s^2 = 1/4):
    a Gaussian's variance grows by 2 D t per axis."""
    n, D, sig0 = 32, 0.05, 2.5
    s = O.OracleSolver(19, (n, n, n), prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=0.05)
    k = s.scalar_enable(D, d3q7)
    x, y, z = cases.coords((n, n, n))
    r2 = (x - 15.5) ** 2 + (y - 15.5) ** 2 + (z - 15.5) ** 2
    s.scalar_set(k, np.exp(-r2 / (2 * sig0 ** 2)))
    T = 200
    s.step(T)
    C = s.scalar_get(k)
    var_x = (C * (x - 15.5) ** 2).sum() / C.sum()
    assert abs(var_x / (sig0 ** 2 + 2 * D * T) - 1) < 0.01


# ---- QOIs (qoi.rs:951-1015) -------------------------------------------------------

def test_power_qois_reference_cases():
    z = O.power_qois(0.0, 2 * np.pi, 1000.0, 0.1, 0.01)
    assert z["power"] == 0.0 and z["np"] == 0.0
    q = O.power_qois(2.0, 2 * np.pi * 3.0, 1000.0, 0.5, 2.0)
    assert abs(q["power"] - 2 * np.pi * 6.0) < 1e-12
    assert abs(q["p_over_v"] - q["power"] / 2.0) < 1e-12
    assert abs(q["np"] - q["power"] / (1000.0 * 27.0 * 0.5 ** 5)) < 1e-15
    assert O.power_qois(-2.0, 2 * np.pi, 1000.0, 0.5, 1.0)["power"] < 0
    with pytest.raises(ValueError):
        O.power_qois(1.0, 2 * np.pi, -1.0, 0.5, 1.0)


def test_mixing_time_and_cv_reference_cases():
    m = O.mixing_time([0.0, 1.0, 2.0, 3.0], [1.0, 0.2, 0.05, 0.009])
    assert m == {"cv0": 1.0, "t95": 2.0, "t99": 3.0}
    assert O.scalar_cv(np.ones(8), np.ones(8, np.uint8)) == 0.0
    assert O.mixing_time([0.0, 1.0], [0.0, 0.0]) is None
    v = np.array([1.0, 2.0, 3.0, 100.0])
    inc = np.array([1, 1, 1, 0], np.uint8)
    assert abs(O.scalar_cv(v, inc) - np.std([1.0, 2.0, 3.0]) / 2.0) < 1e-15


def test_compartment_cv_partitions_fluid():
    n = 12
    solid = O.tank_mask((n, n, n), 1.0, 10.0, 10.0)
    v = np.linspace(1.0, 2.0, n ** 3)
    out = O.compartment_cv((n, n, n), v, solid, imp_z=4.0)
    assert sum(cnt for _, cnt in out.values()) == (solid == 0).sum()


@pytest.mark.parametrize("imp_z", [None, 4.0, 9.5])
def test_compartment_cv_matches_brute_force(imp_z):
    """compartment_cv (qoi.rs:699-739) against numpy on a non-cubic grid:
    top where (z + 0.5)/nz >= 2/3, near_impeller within nz/6 of the impeller
    height, bulk otherwise; CV = population std / |mean| per compartment."""
    dims = (11, 9, 14)
    nx, ny, nz = dims
    rng = np.random.default_rng(5)
    solid = (rng.random(nx * ny * nz) < 0.2).astype(np.uint8)
    v = 1.0 + rng.random(nx * ny * nz)
    out = O.compartment_cv(dims, v, solid, imp_z=imp_z)
    z = np.repeat(np.arange(nz), nx * ny) + 0.5
    fl = solid == 0
    top = fl & (z / nz >= 2.0 / 3.0)
    near = fl & ~top & (np.abs(z - imp_z) <= nz / 6.0) if imp_z is not None else np.zeros_like(fl)
    bulk = fl & ~top & ~near
    for name, m in (("top", top), ("bulk", bulk), ("near_impeller", near)):
        cv, cnt = out[name]
        assert cnt == m.sum()
        if cnt == 0:
            assert cv is None
        else:
            assert cv == pytest.approx(np.std(v[m]) / abs(np.mean(v[m])), rel=1e-12)


# ---- restated reference tests: moving walls, advection, IBM sweeps ------------------

@pytest.mark.parametrize("Q,coll", [(19, "trt"), (27, "trt"), (27, "cm")])
def test_moving_wall_plane_couette_is_exact(Q, coll):
    """Half-way bounce-back with a moving wall (stream_row, kernels.rs:595-603):
    plane Couette between a resting wall and a lid moving in +x has the exact
    linear profile u(y) = U (y - 1/2) / (ny - 2) at steady state (walls on
    the half-way planes).  The sign and the factor 6 w rho c.u of the wall
    term both show in the profile."""
    nx, ny, nz, U = 3, 10, 1, 0.01
    u = np.zeros((6, 3))
    u[O.YPOS] = [U, 0.0, 0.0]
    solid, wall = O.build_wall_rims((nx, ny, nz), [0, 0, 1, 1, 0, 0], u)
    s = O.OracleSolver(Q, (nx, ny, nz), prec=8, periodic=(1, 0, 1), collision=cases.COLL[coll], nu=1.0 / 6.0,
                       solid=solid, wall_u=wall)
    s.step(3000)
    _, ux, uy, uz = (m.reshape(nz, ny, nx)[0] for m in s.moments())
    exact = U * (np.arange(1, ny - 1) - 0.5) / (ny - 2)
    # TRT: exact to round-off; central moment: its order-2 rate carries
    # omega_s (1 - 0.16 |u|^2) (kernels.rs:481-490), a viscosity that varies
    # with |u| by O(u^2) and bends the profile by a few 1e-6 of U
    tol = 1e-10 if coll != "cm" else 1e-4
    for col in range(nx):
        assert np.abs(ux[1:ny - 1, col] - exact).max() < tol * U
    assert np.abs(uy[1:ny - 1]).max() < 1e-13 and np.abs(uz[1:ny - 1]).max() < 1e-13


def _centroid_drift(Q_scalar_d3q7, u0):
    """solver.rs:5996-6033 restated with its drift made exact: a slab of
    tracer in a uniform periodic stream; the tracer starts at its
    equilibrium (set_scalar_concentration, solver.rs:2946-2956), so the
    summed first moment of h is u0 * total at every step and the centroid
    moves exactly u0 per step."""
    dims = (40, 40, 40)
    s = O.OracleSolver(19, dims, prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=0.08)
    n = int(np.prod(dims))
    s.init_with(np.ones(n), np.tile(u0, (n, 1)))
    k = s.scalar_enable(0.01, Q_scalar_d3q7)
    x, y, z = cases.coords(dims)
    c0 = ((np.abs(x - 20) <= 1) & (np.abs(y - 20) <= 1) & (np.abs(z - 20) <= 1)).astype(np.float64)
    s.scalar_set(k, c0)
    steps = 10
    s.step(steps)
    c1 = s.scalar_get(k)

    def centroid(c, coord):
        return (c * coord).sum() / c.sum()

    return [(centroid(c1, a) - centroid(c0, a)) / steps for a in (x, y, z)], c1


@pytest.mark.parametrize("d3q7", [False, True])
def test_scalar_advects_with_the_stream(d3q7):
    u0 = np.array([0.04, -0.01, 0.02])
    drift, c1 = _centroid_drift(d3q7, u0)
    np.testing.assert_allclose(drift, u0, rtol=1e-10, atol=1e-14)
    # no tracer crossed a periodic seam in 10 steps (the moments above are unwrapped)
    x, y, z = cases.coords((40, 40, 40))
    far = (np.abs(x - 20) > 12) | (np.abs(y - 20) > 12) | (np.abs(z - 20) > 12)
    assert np.abs(c1[far]).max() < 1e-30


def _circle_2d(center, radius, omega_z, n_markers):
    """RotatingBody::circle_2d (rotating_ibm.rs:55-75): n markers on a circle
    in the z = 0 plane, weight = arc length 2 pi R / n."""
    th = 2 * np.pi * np.arange(n_markers) / n_markers
    pos = np.stack([center[0] + radius * np.cos(th), center[1] + radius * np.sin(th), 0 * th], 1)
    w = np.full(n_markers, 2 * np.pi * radius / n_markers)
    return pos, w, np.array([center[0], center[1], 0.0]), np.array([0.0, 0.0, omega_z])


def _plane_solver(nx, ny, solid=None, wall=None, periodic=(1, 1, 1)):
    """A one-cell-deep periodic-z D3Q19 TRT box, nu = 1/6: with u_z = 0 its
    populations summed over c_z evolve as D2Q9 (the reference tests' 2D
    lattice), and with the 2D marker stencil the reference's 2D IBM bands
    apply."""
    return O.OracleSolver(19, (nx, ny, 1), prec=8, periodic=periodic, collision=O.TRT, nu=1.0 / 6.0, solid=solid,
                          wall_u=wall)


def test_ibm_iterations_reduce_slip_and_conserve_momentum():
    """rotating_ibm.rs:71-106: 96 markers (kernel radius 2: overlapping
    stencils, so the mobility's row sum differs from the diagonal), one
    sweep against sixteen."""
    pos, w, c, om = _circle_2d((24.0, 24.0), 8.0, 0.003, 96)
    single = _plane_solver(48, 48).apply_ibm(pos, w, c, om, max_iter=1, tol=0.0, radius=2, stencil_dim=2)
    multi = _plane_solver(48, 48).apply_ibm(pos, w, c, om, max_iter=16, tol=0.0, radius=2, stencil_dim=2)
    assert multi["iterations"] > 1
    assert multi["slip_max_rel"] < 0.5 * single["slip_max_rel"]
    assert multi["slip_max_rel"] < 6.0e-4
    assert multi["slip_rms_rel"] < 4.0e-4
    assert multi["momentum_error_rel"] < 1.0e-3


def test_ibm_couette_matches_native_moving_wall():
    """rotating_ibm.rs:176-218: a marker line moving at U one cell below a
    resting wall drives nearly the profile the native moving wall drives."""
    nx, ny, top = 48, 34, 0.002
    u = np.zeros((6, 3))
    u[O.YPOS] = [top, 0.0, 0.0]
    solid, wall = O.build_wall_rims((nx, ny, 1), [0, 0, 1, 1, 0, 0], u)
    native = _plane_solver(nx, ny, solid, wall, periodic=(1, 0, 1))
    native.step(3000)
    solid0, wall0 = O.build_wall_rims((nx, ny, 1), [0, 0, 1, 1, 0, 0], np.zeros((6, 3)))
    ibm = _plane_solver(nx, ny, solid0, wall0, periodic=(1, 0, 1))
    yl = float(ny - 3)
    omega = 1.0e-6
    pos = np.stack([np.arange(nx, dtype=float), np.full(nx, yl), np.zeros(nx)], 1)
    c = np.array([0.0, yl + top / omega, 0.0])
    for _ in range(800):
        ibm.clear_force_field()
        last = ibm.apply_ibm(pos, np.ones(nx), c, np.array([0.0, 0.0, omega]), max_iter=1, tol=1.0, radius=1,
                             relax=0.05, stencil_dim=2)
        ibm.step()
    a = native.moments()[1].reshape(ny, nx)[:, nx // 2]
    b = ibm.moments()[1].reshape(ny, nx)[:, nx // 2]
    sl = slice(2, ny - 3)
    l2 = np.sqrt(((b[sl] - a[sl]) ** 2).sum() / (a[sl] ** 2).sum())
    assert last["slip_max_rel"] < 2.0e-3
    assert l2 < 0.75
    assert np.abs(b[sl] - a[sl]).max() / top < 0.65
    # both profiles rise toward the moving boundary (the sign of the drive)
    assert a[ny - 3] > a[ny // 2] > a[3] > 0 and b[ny - 4] > b[ny // 2] > b[3] > 0


def test_ibm_taylor_couette_torque_and_profile():
    """rotating_ibm.rs:222-309: inner cylinder of markers (r_i = 10) turning
    at Omega inside a resting half-way wall (r_o = 30); the torque is within
    10 % of 4 pi nu Omega r_i^2 r_o^2 / (r_o^2 - r_i^2) and the azimuthal
    profile within the reference's bands of A r + B / r."""
    nx = ny = 80
    cx = cy = 40.0
    r_i, r_o, omega = 10.0, 30.0, 0.00015
    x, y, _ = cases.coords((nx, ny, 1))
    r = np.hypot(x - cx, y - cy)
    solid = (r > r_o + 0.5).astype(np.uint8)
    sim = _plane_solver(nx, ny, solid, np.zeros((nx * ny, 3)))
    pos, w, c, om = _circle_2d((cx, cy), r_i, omega, 160)
    for _ in range(800):
        sim.clear_force_field()
        last = sim.apply_ibm(pos, w, c, om, max_iter=4, tol=1.0e-3, radius=2, relax=1.0, stencil_dim=2)
        sim.step()
    _, ux, uy, _ = sim.moments()
    ro = r_o + 1.0
    A = -omega * r_i ** 2 / (ro ** 2 - r_i ** 2)
    B = omega * r_i ** 2 * ro ** 2 / (ro ** 2 - r_i ** 2)
    band = (r >= r_i + 4.0) & (r <= r_o - 5.0)
    ut = (-ux * (y - cy) + uy * (x - cx))[band] / r[band]
    exact = A * r[band] + B / r[band]
    l2 = np.sqrt(((ut - exact) ** 2).sum() / (exact ** 2).sum())
    torque_ref = 4 * np.pi * (1.0 / 6.0) * omega * r_i ** 2 * ro ** 2 / (ro ** 2 - r_i ** 2)
    assert last["slip_max_rel"] < 1.0e-5
    assert last["momentum_error_rel"] < 1.0e-6
    assert abs(abs(last["torque"][2]) - torque_ref) / torque_ref < 0.10
    assert l2 < 7.0e-2
    assert np.abs(ut - exact).max() / (omega * r_i) < 6.0e-2


@pytest.mark.parametrize("force", [(0.0, 0.0, 0.0), (3e-5, -2e-5, 1e-5)])
def test_cm_factorised_matches_the_generic_matrix_path(force):
    """The oracle's factorised D3Q27 central-moment transforms against the
    reference's generic matrix path (kernels.rs:227-392) on the same random
    state with a per-cell force and shear-rate field, over several steps of
    a box with walls: agreement to f64 round-off."""
    dims = (9, 8, 7)
    solid = np.asarray(O.build_wall_rims(dims, [1] * 6, np.zeros((6, 3)))[0])
    rng = np.random.default_rng(11)
    n = int(np.prod(dims))
    ff = np.where((solid == 0)[:, None], rng.normal(0, 1e-5, (n, 3)), 0.0)
    om = 1.0 / (3.0 * (0.02 + 0.01 * rng.random(n)) + 0.5)
    runs = []
    for fact in (True, False):
        O.set_cm_factorised(fact)
        try:
            s = O.OracleSolver(27, dims, prec=8, collision=O.CM, nu=0.02, omega_shear=1.7, solid=solid,
                               force=force)
            _random_state(s, 27, np.random.default_rng(3), amp=3e-3)
            s.set_force_field(ff)
            s.set_omega_field(om)
            s.step(4)
            runs.append(s.get_f())
        finally:
            O.set_cm_factorised(True)
    a, b = runs
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b + O.lattice(27)[1][:, None]).max()
