"""GPU parity of the face patches (FacePatch, params.rs:286-296; the patch
arms of apply_open_faces_impl, backend.rs:602-766) against the oracle on
identical inputs: fp32 populations and moments bit-exact after odd and even
step counts -- a closed face with velocity / pressure / outflow / convective
/ closed patches (zero-velocity lid outside them), patches over an open
base face, patches on faces meeting at an edge, both lattices, body force --
and the refusals of validate_face_patches (solver.rs:690-740)."""
import numpy as np
import pytest

import lbmflow_amd as L
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

NAME = {O.FACE_CLOSED: "closed", O.FACE_VELOCITY: "velocity", O.FACE_PRESSURE: "pressure",
        O.FACE_OUTFLOW: "outflow", O.FACE_CONVECTIVE: "convective"}


def rims(dims, faces):
    nx, ny, nz = dims
    s = np.zeros((nz, ny, nx), np.uint8)
    for f in faces:
        sl = [slice(None)] * 3
        sl[2 - f // 2] = 0 if f % 2 == 0 else dims[f // 2] - 1
        s[tuple(sl)] = 1
    return s.reshape(-1)


def f32v(value):
    return None if value is None else tuple(np.atleast_1d(np.asarray(value, np.float32)).astype(np.float64))


def set_patches(orc, hip, patches):
    p32 = [(f, lo, hi, k, f32v(v)) for f, lo, hi, k, v in patches]
    orc.set_face_patches(p32)
    hip.set_face_patches([(f, lo, hi, NAME[k], v) for f, lo, hi, k, v in p32])


def set_face(orc, hip, face, kind, value=None):
    orc.set_face(face, kind, f32v(value) if value is not None else (0.0,))
    hip.set_face_bc(face, NAME[kind], f32v(value))


def check(orc, hip, solid, what):
    fl = solid == 0
    a = hip.populations()[:, fl].astype(np.float64)
    b = orc.get_f()[:, fl]
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} populations differ, first {bad[:3]}, max {np.abs(a - b).max()}"
    for k, (x, y) in enumerate(zip(hip.moments(), orc.moments())):
        assert np.array_equal(x[fl].astype(np.float64), y[fl]), f"{what}: moment {k}"


def run_pair(orc, hip, solid, counts, what):
    done = 0
    for k in counts:
        orc.step(k)
        hip.run(k)
        done += k
        check(orc, hip, solid, f"{what} step {done}")


@pytest.mark.parametrize("Q", [19, 27])
def test_closed_face_with_patches_bitexact(Q):
    dims = (12, 10, 8)
    solid = rims(dims, [O.XNEG, O.XPOS, O.YNEG, O.ZNEG, O.ZPOS])
    orc, hip = cases.pair(Q, dims, collision="trt", nu=0.05, solid=solid)
    # y+ face, in-face coordinates (x, z)
    set_patches(orc, hip, [(O.YPOS, (3, 2), (5, 4), O.FACE_VELOCITY, (0.01, -0.03, 0.005)),
                           (O.YPOS, (7, 1), (9, 3), O.FACE_PRESSURE, (1.0,)),
                           (O.YPOS, (7, 4), (8, 6), O.FACE_OUTFLOW, None),
                           (O.YPOS, (1, 5), (2, 6), O.FACE_CONVECTIVE, (0.08,)),
                           (O.YPOS, (10, 1), (10, 6), O.FACE_CLOSED, None)])
    run_pair(orc, hip, solid, (1, 1, 6, 13), f"D3Q{Q} closed y+ with patches")


@pytest.mark.parametrize("Q", [19, 27])
def test_patches_over_open_faces_at_an_edge_bitexact(Q):
    """Open base rules with patches on two faces sharing an edge (y+ and
    z+), and a body force (the Zou-He raw-momentum branch)."""
    dims = (12, 10, 8)
    solid = rims(dims, [O.XNEG, O.XPOS, O.YNEG, O.ZNEG])
    orc, hip = cases.pair(Q, dims, collision="trt", nu=0.05, solid=solid, force=(2e-6, -1e-6, 1e-6))
    set_face(orc, hip, O.YPOS, O.FACE_OUTFLOW)
    set_face(orc, hip, O.ZPOS, O.FACE_PRESSURE, (1.0,))
    set_patches(orc, hip, [(O.YPOS, (3, 2), (5, 7), O.FACE_VELOCITY, (0.0, -0.02, 0.0)),
                           (O.ZPOS, (2, 2), (4, 9), O.FACE_VELOCITY, (0.004, 0.0, -0.01)),
                           (O.ZPOS, (6, 6), (11, 9), O.FACE_CONVECTIVE, (0.05,))])
    run_pair(orc, hip, solid, (1, 2, 9), f"D3Q{Q} open edge with patches")


def test_impinging_jet_bitexact():
    """The impinging-jet layout of t18_2_masked_face.rs:73-103 (BGK, top
    face closed around a velocity patch and a pressure annulus)."""
    dims = (32, 32, 24)
    solid = rims(dims, [O.XNEG, O.XPOS, O.YNEG, O.YPOS, O.ZNEG])
    orc, hip = cases.pair(19, dims, collision="bgk", nu=0.04, solid=solid)
    jet, outlet = (O.FACE_VELOCITY, (0.0, 0.0, -0.05)), (O.FACE_PRESSURE, (1.0,))
    set_patches(orc, hip, [(O.ZPOS, (14, 14), (17, 17)) + jet, (O.ZPOS, (11, 11), (20, 13)) + outlet,
                           (O.ZPOS, (11, 18), (20, 20)) + outlet, (O.ZPOS, (11, 14), (13, 17)) + outlet,
                           (O.ZPOS, (18, 14), (20, 17)) + outlet])
    run_pair(orc, hip, solid, (5, 40), "impinging jet")


def test_patches_in_a_captured_span_match_eager():
    """A rotating body under a patched lid: the span graph (patch table in
    its kernel arguments) equals per-step stepping."""
    dims = (16, 12, 12)
    solid = rims(dims, [O.XNEG, O.XPOS, O.YNEG, O.YPOS, O.ZNEG])
    th = np.linspace(0, 2 * np.pi, 20, endpoint=False)
    pos = np.stack([8 + 2.5 * np.cos(th), 6 + 2.5 * np.sin(th), np.full_like(th, 5.0)], 1)
    runs = []
    for graph in (False, True):
        s = L.Solver(19, dims, nu=0.04, collision="bgk", solid=solid)
        s.set_face_patches([("z+", (6, 4), (9, 7), "velocity", (0.0, 0.0, -0.03)),
                            ("z+", (1, 1), (3, 10), "pressure", (1.0,))])
        s.set_rotating_body(pos, np.ones(len(th)), np.array([8.0, 6.0, 5.0]), (0.0, 0.0, 0.01))
        tq, _ = s.run_with_ibm(5, graph=graph)
        tq2, _ = s.run_with_ibm(6, graph=graph)
        runs.append((s.populations(), tq, tq2))
        s.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])


def test_refusals():
    dims = (12, 10, 8)
    s = L.Solver(19, dims, nu=0.05, periodic=(True, False, False))
    assert s.ctx.capabilities() & 0x4000  # LBM_CAP_LOCALIZED
    bad = [
        [("z+", (0, 0), (12, 9), "outflow", None)],                       # hi beyond the face
        [("z+", (5, 0), (4, 3), "outflow", None)],                        # inverted
        [("z+", (4, 4), (8, 8), "outflow", None), ("z+", (8, 6), (11, 9), "pressure", (1.0,))],  # overlap
        [("z+", (6, 6), (9, 9), "velocity", (0.0, 0.0, -0.3000001))],    # above MAX_SPEED
        [("z+", (6, 6), (9, 9), "pressure", (0.0,))],                     # density not positive
        [("x+", (1, 1), (2, 2), "closed", None)],                         # periodic axis
        [("z+", (k, 0), (k, 0), "outflow", None) for k in range(12)] + [("z-", (0, 0), (0, 0), "outflow", None)] * 5,
    ]
    for patches in bad:
        with pytest.raises(L.LbmError):
            s.set_face_patches(patches)
    with pytest.raises(L.LbmError):
        s.set_face_bc("z+", "velocity", (0.0, 0.0, -0.31))
    # the same rect on two faces is fine, and a closed face covered by
    # patches needs no wall rim
    s.set_face_patches([("z+", (4, 4), (8, 8), "outflow", None), ("z-", (4, 4), (8, 8), "outflow", None),
                        ("y-", (0, 0), (0, 0), "closed", None), ("y+", (0, 0), (0, 0), "closed", None)])
    s.run(2)
    s.set_face_patches([])
    with pytest.raises(L.LbmError):  # the faces are uncovered again
        s.run(1)
    s.close()
