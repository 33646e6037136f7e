"""Pins of the oracle's face patches (FacePatch, params.rs:286-296; the patch
arms of apply_open_faces_impl, backend.rs:602-766), anchored on the
reference's own acceptance tests (tests/t18_2_masked_face.rs,
tests/face_patch_smoke.rs) and on what the rules fix exactly:

  * a full-face velocity patch on a closed face equals the same velocity as
    the face's base rule (t18_2_full_face_patch_matches_base_face_bc, band
    1e-14; here bit-identical);
  * the impinging jet through a top face closed around a velocity patch and
    a pressure annulus conserves mass to 5e-9 per 200 steps after 2400 steps
    and forms a radial wall jet on the floor with a stagnation minimum on the
    axis, a positive off-axis peak above 1e-5 and monotone decay
    (t18_2_impinging_jet_same_top_face_conserves_mass_and_has_radial_wall_jet);
  * a velocity patch leaves its cells at exactly the prescribed velocity, the
    closed base cells of that face at exactly zero velocity, a pressure patch
    its cells at exactly its density, and a patch over an open base face
    overrides the base rule on its rect only;
  * out-of-bounds and overlapping rects are refused
    (t18_2_rejects_patch_out_of_face_bounds / _overlapping_patches).
"""
import numpy as np
import pytest

from oracle import oracle as O


def rims(dims, faces):
    """build_wall_rims (solver.rs:1357-1400): every cell of each listed face
    plane is solid."""
    nx, ny, nz = dims
    s = np.zeros((nz, ny, nx), np.uint8)
    for f in faces:
        sl = [slice(None)] * 3
        ax = 2 - f // 2  # array axes are (z, y, x)
        sl[ax] = 0 if f % 2 == 0 else dims[f // 2] - 1
        s[tuple(sl)] = 1
    return s.reshape(-1)


def fields(o, dims):
    nx, ny, nz = dims
    return [v.reshape(nz, ny, nx) for v in o.moments()]


def test_full_face_patch_matches_base_face_rule():
    dims = (18, 12, 10)
    solid = rims(dims, [O.YNEG, O.YPOS, O.ZNEG, O.ZPOS])
    runs = []
    for patched in (False, True):
        o = O.OracleSolver(19, dims, prec=8, collision=O.BGK, nu=0.04, solid=solid)
        o.set_face(O.XPOS, O.FACE_PRESSURE, (1.0,))
        if patched:
            o.set_face_patches([(O.XNEG, (0, 0), (dims[1] - 1, dims[2] - 1), O.FACE_VELOCITY, (0.035, 0.0, 0.0))])
        else:
            o.set_face(O.XNEG, O.FACE_VELOCITY, (0.035, 0.0, 0.0))
        o.step(80)
        runs.append((o.get_f(), o.moments()))
    assert np.array_equal(runs[0][0], runs[1][0])
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)


def impinging_jet():
    dims = (32, 32, 24)
    solid = rims(dims, [O.XNEG, O.XPOS, O.YNEG, O.YPOS, O.ZNEG])
    o = O.OracleSolver(19, dims, prec=8, collision=O.BGK, nu=0.04, solid=solid)
    jet, outlet = (O.FACE_VELOCITY, (0.0, 0.0, -0.05)), (O.FACE_PRESSURE, (1.0,))
    o.set_face_patches([(O.ZPOS, (14, 14), (17, 17)) + jet,
                        (O.ZPOS, (11, 11), (20, 13)) + outlet,
                        (O.ZPOS, (11, 18), (20, 20)) + outlet,
                        (O.ZPOS, (11, 14), (13, 17)) + outlet,
                        (O.ZPOS, (18, 14), (20, 17)) + outlet])
    return o, dims, solid


def test_impinging_jet_conserves_mass_and_has_radial_wall_jet():
    o, dims, solid = impinging_jet()
    nx, ny, nz = dims
    o.step(2400)
    rho = fields(o, dims)[0]
    fl = solid.reshape(nz, ny, nx) == 0
    m0 = rho[fl].sum()
    o.step(200)
    rho, ux, uy, _ = fields(o, dims)
    m1 = rho[fl].sum()
    assert abs(m1 - m0) / abs(m0) <= 5e-9
    cx, cy = (nx - 1) * 0.5, (ny - 1) * 0.5
    profile = []
    radii = [0, 2, 4, 6, 8, 10, 12]
    for r in radii:
        vals = []
        for y in range(1, ny - 1):
            for x in range(1, nx - 1):
                dx, dy = x - cx, y - cy
                rr = np.hypot(dx, dy)
                if abs(rr - r) <= 0.5 or (r == 0 and rr <= 1.0):
                    vals.append((ux[1, y, x] * dx + uy[1, y, x] * dy) / max(rr, 1e-12))
        profile.append(np.mean(vals))
    peak_idx = 1 + int(np.argmax(profile[1:]))
    peak = profile[peak_idx]
    assert abs(profile[0]) <= 0.25 * abs(peak), profile
    assert peak > 1e-5, profile
    for i in range(peak_idx + 1, len(profile)):
        assert profile[i] <= profile[i - 1] + 5e-5, profile


@pytest.mark.parametrize("Q", [19, 27])
def test_patch_rules_hold_exactly_on_their_cells(Q):
    dims = (12, 10, 8)
    nx, ny, nz = dims
    solid = rims(dims, [O.XNEG, O.XPOS, O.YNEG, O.ZNEG])
    o = O.OracleSolver(Q, dims, prec=8, collision=O.TRT, nu=0.05, solid=solid)
    u_jet = (0.004, 0.0, 0.0)
    # y+ face: in-face coordinates (x, z)
    o.set_face_patches([(O.YPOS, (3, 2), (5, 4), O.FACE_VELOCITY, (0.01, -0.03, 0.005)),
                        (O.YPOS, (7, 2), (9, 5), O.FACE_PRESSURE, (1.0,)),
                        (O.ZPOS, (2, 2), (4, 3), O.FACE_VELOCITY, u_jet)])
    o.set_face(O.ZPOS, O.FACE_OUTFLOW)
    fl = solid.reshape(nz, ny, nx) == 0
    for _ in range(4):
        o.step(9)
        rho, ux, uy, uz = fields(o, dims)
        top = np.s_[2:5, ny - 1, 3:6]
        np.testing.assert_allclose(ux[top], 0.01, rtol=0, atol=1e-15)
        np.testing.assert_allclose(uy[top], -0.03, rtol=0, atol=1e-15)
        np.testing.assert_allclose(uz[top], 0.005, rtol=0, atol=1e-15)
        np.testing.assert_allclose(rho[2:6, ny - 1, 7:10], 1.0, rtol=0, atol=1e-14)
        # the closed base of y+ outside the patches: an impermeable, resting lid
        lid = fl[:, ny - 1, :].copy()
        lid[2:5, 3:6] = lid[2:6, 7:10] = False
        lid[nz - 1] = False  # the y+/z+ edge: the z+ rules run after y+ (Face::ALL order)
        assert lid.sum() > 0
        for v in (ux, uy, uz):
            np.testing.assert_allclose(v[:, ny - 1, :][lid], 0.0, rtol=0, atol=1e-15)
        # z+ face: patch cells take the jet velocity, the open base elsewhere
        # is zero-gradient outflow (not pinned to any velocity)
        for v, c in zip((ux, uy, uz), u_jet):
            np.testing.assert_allclose(v[nz - 1, 2:4, 2:5], c, rtol=0, atol=1e-15)
    assert np.abs(uz[nz - 1][fl[nz - 1]]).max() > 1e-6


def test_bad_patches_are_refused():
    dims = (12, 10, 8)
    o = O.OracleSolver(19, dims, prec=8, collision=O.BGK, nu=0.04)
    with pytest.raises(ValueError):  # hi beyond the face (x extent 12)
        o.set_face_patches([(O.ZPOS, (0, 0), (12, 9), O.FACE_OUTFLOW, None)])
    with pytest.raises(ValueError):  # inverted
        o.set_face_patches([(O.ZPOS, (5, 0), (4, 3), O.FACE_OUTFLOW, None)])
    with pytest.raises(ValueError):  # overlap on one face
        o.set_face_patches([(O.ZPOS, (4, 4), (8, 8), O.FACE_OUTFLOW, None),
                            (O.ZPOS, (8, 6), (11, 9), O.FACE_PRESSURE, (1.0,))])
    o.set_face_patches([(O.ZPOS, (4, 4), (8, 8), O.FACE_OUTFLOW, None),
                        (O.ZNEG, (4, 4), (8, 8), O.FACE_OUTFLOW, None)])  # same rect, other face
    p = O.OracleSolver(19, dims, prec=8, periodic=(0, 0, 1), collision=O.BGK, nu=0.04)
    with pytest.raises(ValueError):
        p.set_face_patches([(O.ZPOS, (1, 1), (2, 2), O.FACE_CLOSED, None)])
