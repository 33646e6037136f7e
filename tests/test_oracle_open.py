"""Pins of the oracle's open faces (kernels.rs:849-1299, backend.rs:602-705):
facts the boundary rules guarantee, independent of how they are written.

  * Zou-He velocity faces leave the face cells with exactly the prescribed
    velocity, pressure faces with the prescribed density and no tangential
    velocity -- D3Q19 and D3Q27, with and without a body force (the Guo
    half-force correction of the raw momentum);
  * a uniform flow entering through a velocity face and leaving through a
    pressure or outflow face is a fixed point of the whole step;
  * convective faces pin the face density to the inner neighbour's;
  * a velocity-driven channel carries the same mass flux through every
    section and develops the plane-Poiseuille parabola.
"""
import numpy as np
import pytest

from oracle import oracle as O


def channel(Q, dims, nu=0.1, force=(0.0, 0.0, 0.0), walls=True, periodic=(0, 0, 1), prec=8):
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    if walls:
        solid[:, 0, :] = 1
        solid[:, ny - 1, :] = 1
    o = O.OracleSolver(Q, dims, prec=prec, periodic=periodic, collision=O.TRT, nu=nu, solid=solid.reshape(-1),
                       force=force)
    return o, solid


def face_moments(o, dims, x):
    nx, ny, nz = dims
    r, a, b, c = o.moments()
    sl = lambda v: v.reshape(nz, ny, nx)[:, :, x]
    return sl(r), sl(a), sl(b), sl(c)


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("force", [(0.0, 0.0, 0.0), (2e-6, -1e-6, 5e-7)])
def test_velocity_face_imposes_the_velocity(Q, force):
    dims = (12, 8, 6)
    o, solid = channel(Q, dims, force=force, walls=False, periodic=(0, 1, 1))
    u_in = (0.02, 0.004, -0.003)
    o.set_face(O.XNEG, O.FACE_VELOCITY, u_in)
    o.set_face(O.XPOS, O.FACE_PRESSURE, (1.0,))
    for _ in range(5):
        o.step(7)
        r, a, b, c = face_moments(o, dims, 0)
        np.testing.assert_allclose(a, u_in[0], rtol=0, atol=1e-15)
        np.testing.assert_allclose(b, u_in[1], rtol=0, atol=1e-15)
        np.testing.assert_allclose(c, u_in[2], rtol=0, atol=1e-15)
        r, a, b, c = face_moments(o, dims, dims[0] - 1)
        np.testing.assert_allclose(r, 1.0, rtol=0, atol=1e-14)
        np.testing.assert_allclose(b, 0.0, rtol=0, atol=1e-15)
        np.testing.assert_allclose(c, 0.0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("outlet", [O.FACE_PRESSURE, O.FACE_OUTFLOW])
def test_uniform_flow_is_a_fixed_point(Q, outlet):
    dims = (10, 5, 4)
    o, _ = channel(Q, dims, walls=False, periodic=(0, 1, 1))
    u0 = (0.03, 0.0, 0.0)
    n = dims[0] * dims[1] * dims[2]
    o.init_with(np.ones(n), np.tile(u0, (n, 1)))
    f0 = o.get_f()
    o.set_face(O.XNEG, O.FACE_VELOCITY, u0)
    o.set_face(O.XPOS, outlet, (1.0,))
    o.step(40)
    assert np.abs(o.get_f() - f0).max() < 1e-15


@pytest.mark.parametrize("Q", [19, 27])
def test_convective_face_pins_density_to_the_inner_cell(Q):
    dims = (16, 8, 4)
    o, solid = channel(Q, dims)
    o.set_face(O.XNEG, O.FACE_VELOCITY, (0.02, 0.0, 0.0))
    o.set_face(O.XPOS, O.FACE_CONVECTIVE, (0.05,))
    fl = solid[:, :, -1] == 0
    o.step(40)  # let the inlet disturbance reach the outlet
    for _ in range(6):
        o.step(5)
        r_face = face_moments(o, dims, dims[0] - 1)[0]
        r_in = face_moments(o, dims, dims[0] - 2)[0]
        np.testing.assert_allclose(r_face[fl], r_in[fl], rtol=0, atol=1e-14)
        assert np.abs(r_face[fl] - 1.0).max() > 1e-9  # the density does vary


@pytest.mark.parametrize("Q", [19, 27])
def test_developed_channel_flow(Q):
    """Velocity-driven channel between half-way walls after 4000 steps: every
    x-section carries the inlet flux (to the remaining transient), and
    mid-channel the profile has the plane-Poiseuille shape."""
    dims = (40, 10, 2)
    nx, ny, nz = dims
    u_in = 0.01
    o, solid = channel(Q, dims, nu=0.2)
    o.set_face(O.XNEG, O.FACE_VELOCITY, (u_in, 0.0, 0.0))
    o.set_face(O.XPOS, O.FACE_PRESSURE, (1.0,))
    o.step(4000)
    r, a, _, _ = o.moments()
    r = r.reshape(nz, ny, nx)[0]
    a = a.reshape(nz, ny, nx)[0]
    flux = (r[1:-1] * a[1:-1]).sum(axis=0)
    np.testing.assert_allclose(flux, flux[0], rtol=2e-4)
    H = ny - 2
    y = np.arange(1, ny - 1) + 0.5 - 1.0  # cell centres from the lower half-way wall
    shape = y * (H - y)
    mid = a[1:-1, nx // 2]
    # the parabola's shape (normalised by the mean over the cell centres)
    assert np.abs(mid / mid.mean() - shape / shape.mean()).max() < 1e-4


def test_open_face_on_periodic_axis_is_refused():
    o, _ = channel(19, (6, 6, 6), periodic=(1, 0, 0))
    with pytest.raises(ValueError):
        o.set_face(O.XNEG, O.FACE_OUTFLOW)


def duct_profile(n1, n2, umax=0.03):
    """u_x(c1, c2) = umax f(c1) g(c2) on an x face, face order c2 * n1 + c1."""
    c1 = np.arange(n1) + 0.5
    c2 = np.arange(n2) + 0.5
    f = 4 * c1 * (n1 - c1) / n1**2
    g = 4 * c2 * (n2 - c2) / n2**2
    v = np.zeros((n2, n1, 3))
    v[:, :, 0] = umax * np.outer(g, f)
    v[:, :, 1] = 0.002 * np.outer(g, np.ones(n1))
    return v.reshape(-1, 3)


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("force", [(0.0, 0.0, 0.0), (2e-6, -1e-6, 5e-7)])
def test_inlet_profile_is_imposed_per_cell(Q, force):
    """set_inlet_profile (solver.rs:3322-3350; kernels.rs:1021-1034,
    1164-1180): each face cell of the velocity face takes its own u."""
    dims = (12, 8, 6)
    nx, ny, nz = dims
    o, _ = channel(Q, dims, force=force, walls=False, periodic=(0, 1, 1))
    prof = duct_profile(ny, nz)
    o.set_face(O.XNEG, O.FACE_VELOCITY, (0.01, 0.0, 0.0))
    o.set_inlet_profile(O.XNEG, prof)
    o.set_face(O.XPOS, O.FACE_PRESSURE, (1.0,))
    want = prof.reshape(nz, ny, 3)
    for _ in range(3):
        o.step(7)
        _, a, b, c = face_moments(o, dims, 0)
        for k, m in enumerate((a, b, c)):
            np.testing.assert_allclose(m, want[:, :, k], rtol=0, atol=1e-15)


def test_constant_profile_equals_the_plain_face():
    dims = (10, 6, 5)
    runs = []
    for with_profile in (False, True):
        o, _ = channel(19, dims, walls=False, periodic=(0, 1, 1))
        o.set_face(O.XNEG, O.FACE_VELOCITY, (0.02, 0.001, -0.003))
        if with_profile:
            o.set_inlet_profile(O.XNEG, np.tile((0.02, 0.001, -0.003), (dims[1] * dims[2], 1)))
        o.set_face(O.XPOS, O.FACE_OUTFLOW)
        o.step(15)
        runs.append(o.get_f())
    assert np.array_equal(runs[0], runs[1])


def test_inlet_profile_needs_a_velocity_face():
    o, _ = channel(19, (8, 6, 4), walls=False, periodic=(0, 1, 1))
    o.set_face(O.XNEG, O.FACE_PRESSURE, (1.0,))
    with pytest.raises(ValueError):
        o.set_inlet_profile(O.XNEG, np.zeros((24, 3)))
