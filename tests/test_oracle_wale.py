"""Pins of the oracle's WALE closure (les.rs:120-211) and of the velocity
gradient it reads (strain_rate_at, solver.rs:5393-5447;
gather_velocity_gradient, solver.rs:5490-5581).

The reference's tests hold no WALE vectors, so the pins are facts the
mathematics fixes:
  * a linear velocity field whose populations carry its strain
    (fourth-order lattice isotropy) has the gradient G everywhere, also at
    non-periodic domain faces (first-order differences are exact there);
  * half-way walls: the one-sided wall difference is exact for a quadratic
    profile through the wall velocity at distance 1/2;
  * WALE closed forms: nu_t = 0 in pure shear, Cw^2 (2/3)^(1/4) |w| in
    solid-body rotation, Cw^2 |a| 6^1.5 / (6^2.5 + 6^1.25) in axisymmetric
    strain;
  * body-force Poiseuille flow (exact for TRT with magic 3/16): S_xy is
    the analytic du/dy / 2;
  * the Guo-force term of the stress: a uniformly accelerated flow has zero
    strain, which the corrected stress reproduces to round-off while the raw
    stress does not.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import wale_cases as W


def _oracle(Q, dims, periodic=(0, 0, 0), nu=0.1, prec=8, solid=None, wall_u=None, force=(0.0, 0.0, 0.0),
            collision=O.TRT):
    return O.OracleSolver(Q, dims, prec=prec, periodic=periodic, collision=collision, nu=nu, solid=solid,
                          wall_u=wall_u, force=force)


def _tau(nu):
    return 1.0 / O.omegas(O.TRT, nu)[0]


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 1)])
def test_linear_field_gradient_is_exact(Q, periodic):
    rng = np.random.default_rng(3)
    G = 1e-3 * rng.standard_normal((3, 3))
    if periodic != (0, 0, 0):  # a periodic axis needs u independent of it
        G[:, 0] = 0.0
        G[:, 2] = 0.0
    dims = (6, 5, 7)
    nu = 0.08
    f, rho, u = W.linear_state(Q, dims, G, _tau(nu))
    o = _oracle(Q, dims, periodic, nu)
    o.set_f(f)
    o.update_moments()
    r, a, b, c = o.moments()
    np.testing.assert_allclose(r, rho, rtol=0, atol=1e-14)
    np.testing.assert_allclose(np.stack([a, b, c], 1), u, rtol=0, atol=1e-14)
    S = o.strain()
    Gs = 0.5 * (G + G.T)
    np.testing.assert_allclose(S[:, [0, 1, 2]], np.broadcast_to(np.diag(Gs), (S.shape[0], 3)), atol=1e-13)
    np.testing.assert_allclose(S[:, 3], Gs[0, 1], atol=1e-13)
    np.testing.assert_allclose(S[:, 4], Gs[0, 2], atol=1e-13)
    np.testing.assert_allclose(S[:, 5], Gs[1, 2], atol=1e-13)
    g = o.velocity_gradient()
    np.testing.assert_allclose(g, np.broadcast_to(G, g.shape), rtol=0, atol=1e-13)


@pytest.mark.parametrize("with_strain", [False, True])
def test_half_way_wall_difference_is_exact_for_quadratics(with_strain):
    Q, dims, nu = 19, (4, 11, 3), 0.1
    U, A = 2e-3, 3e-5
    f, solid, wall_u, dudy = W.channel_state(Q, dims, U, A, _tau(nu), with_strain)
    o = _oracle(Q, dims, (1, 0, 1), nu, solid=solid, wall_u=wall_u)
    o.set_f(f)
    o.update_moments()
    g = o.velocity_gradient()
    fl = solid == 0
    expect01 = dudy if with_strain else 0.5 * dudy
    expect10 = 0.0 if with_strain else -0.5 * dudy
    np.testing.assert_allclose(g[fl, 0, 1], expect01[fl], rtol=0, atol=1e-14)
    np.testing.assert_allclose(g[fl, 1, 0], np.broadcast_to(expect10, fl.shape)[fl], rtol=0, atol=1e-14)
    rest = np.ones((3, 3), bool)
    rest[0, 1] = rest[1, 0] = False
    assert np.abs(g[fl][:, rest]).max() < 1e-14
    assert np.abs(g[~fl]).max() == 0.0
    # the wall-adjacent rows are where the one-sided form acts
    y = W.coords(dims)[:, 1]
    assert fl[y == 1].all() and fl[y == dims[1] - 2].all()


def _uniform_gradient_wale(G, nu=0.08, tau_eff_max=None, prec=8):
    dims = (5, 4, 6)
    f, _, _ = W.linear_state(19, dims, G, _tau(nu), rho_amp=0.0)
    o = _oracle(19, dims, (0, 0, 0), nu, prec=prec)
    o.set_f(f)
    o.update_moments()
    return o.wale(tau_eff_max=tau_eff_max)


def test_wale_vanishes_in_pure_shear():
    om, nut, k, mx = _uniform_gradient_wale(np.array([[0, 4e-3, 0], [0, 0, 0], [0, 0, 0]], float))
    assert np.abs(nut).max() < 1e-18 and k == 0
    np.testing.assert_allclose(om, 1.0 / (3.0 * 0.08 + 0.5), rtol=1e-15)


def test_wale_solid_body_rotation_closed_form():
    w = 3e-3
    nu = 0.08
    om, nut, k, mx = _uniform_gradient_wale(np.array([[0, -w, 0], [w, 0, 0], [0, 0, 0]], float), nu)
    expect = O.WALE_CW ** 2 * (2.0 / 3.0) ** 0.25 * w
    np.testing.assert_allclose(nut, expect, rtol=1e-12)
    np.testing.assert_allclose(om, 1.0 / (3.0 * (nu + expect) + 0.5), rtol=1e-12)
    assert mx == pytest.approx(expect, rel=1e-12)


def test_wale_axisymmetric_strain_closed_form():
    a = 2e-3
    om, nut, k, mx = _uniform_gradient_wale(np.diag([a, a, -2 * a]))
    expect = O.WALE_CW ** 2 * a * 6 ** 1.5 / (6 ** 2.5 + 6 ** 1.25)
    np.testing.assert_allclose(nut, expect, rtol=1e-12)


def test_wale_general_gradient_matches_matrix_definition():
    rng = np.random.default_rng(11)
    G = 2e-3 * rng.standard_normal((3, 3))
    om, nut, k, mx = _uniform_gradient_wale(G)
    np.testing.assert_allclose(nut, W.wale_closed_form(G), rtol=1e-11)


def test_wale_clip_bounds_tau_eff():
    w, nu = 3e-3, 0.08
    G = np.array([[0, -w, 0], [w, 0, 0], [0, 0, 0]], float)
    raw = O.WALE_CW ** 2 * (2.0 / 3.0) ** 0.25 * w
    tau0 = 0.5 + 3 * nu
    tmax = tau0 + 1.5 * raw  # nu_t_max = raw / 2
    om, nut, k, mx = _uniform_gradient_wale(G, nu, tau_eff_max=tmax)
    assert k == om.size
    np.testing.assert_allclose(nut, raw / 2, rtol=1e-12)
    np.testing.assert_allclose(om, 1.0 / tmax, rtol=1e-12)
    assert mx == pytest.approx(raw, rel=1e-12)


def test_strain_in_body_force_poiseuille():
    """Steady TRT Poiseuille (magic 3/16: exact parabola with half-way walls):
    u(y) = F/(2 nu) (y - 1/2)(ny - 3/2 - y), so S_xy = du/dy / 2."""
    nu, F = 0.5, 1e-5
    dims = (2, 10, 2)
    ny = dims[1]
    solid = np.zeros(dims[::-1], np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    o = _oracle(19, dims, (1, 0, 1), nu, solid=solid.reshape(-1), force=(F, 0.0, 0.0))
    o.step(600)
    S = o.strain()
    y = W.coords(dims)[:, 1]
    fl = solid.reshape(-1) == 0
    dudy = F / (2 * nu) * (ny - 1.0 - 2.0 * y)
    np.testing.assert_allclose(S[fl, 3], 0.5 * dudy[fl], rtol=0, atol=1e-12)
    assert np.abs(S[~fl]).max() == 0.0


@pytest.mark.parametrize("collision", [O.TRT, O.BGK])
def test_force_term_removes_the_stress_of_uniform_acceleration(collision):
    """A uniform body force in a periodic box accelerates a uniform flow:
    the true strain is zero.  The pre-collision stress alone carries
    -(u F + F u)/2 per step of forcing (about 1e-8 here); with the Guo term
    of strain_rate_at every component cancels to round-off."""
    nu = 0.1
    F = (1e-5, 2e-5, -1e-5)
    o = _oracle(19, (2, 2, 2), (1, 1, 1), nu, force=F, collision=collision)
    o.step(50)
    S = o.strain()
    _, a, b, c = o.moments()
    u = np.array([a[0], b[0], c[0]])
    assert np.abs(u - 50.5 * np.array(F)).max() < 1e-15
    assert np.abs(S).max() < 1e-18
    # the term it removes is not small
    assert np.abs(u[1] * F[1]) / (2 / 3 * _tau(nu)) > 1e-9
