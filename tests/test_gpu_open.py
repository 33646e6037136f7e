"""GPU parity of the open faces (kernels.rs:849-1299, backend.rs:602-705)
against the oracle on identical inputs.

Bars: TRT/BGK fp32 populations and moments bit-exact after odd and even
step counts (the face rules keep the reference's operand order, and the
stale-slot stash reproduces the pull stream's skipped slots exactly);
central-moment D3Q27 within 1e-6; 16-bit storage tracks fp32; a captured
span equals the eager one bit for bit.
"""
import numpy as np
import pytest

import lbmflow_amd as L
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

KIND = {O.FACE_VELOCITY: "velocity", O.FACE_PRESSURE: "pressure", O.FACE_OUTFLOW: "outflow",
        O.FACE_CONVECTIVE: "convective"}


def set_faces(orc, hip, faces):
    for face, kind, value in faces:
        v = tuple(np.asarray(value, np.float32).astype(np.float64)) if value is not None else (0.0,)
        orc.set_face(face, kind, v)
        hip.set_face_bc(face, KIND[kind], v)


def channel_solid(dims, walls_y=True):
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    if walls_y:
        solid[:, 0, :] = 1
        solid[:, ny - 1, :] = 1
    return solid.reshape(-1)


def assert_state(orc, hip, solid, what, exact=True, tol=0.0):
    fl = solid == 0
    a = hip.populations()[:, fl].astype(np.float64)
    b = orc.get_f()[:, fl]
    if exact:
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{what}: {len(bad)} populations differ, first {bad[:3]}, max {np.abs(a - b).max()}"
        for k, (x, y) in enumerate(zip(hip.moments(), orc.moments())):
            assert np.array_equal(x[fl].astype(np.float64), y[fl]), f"{what}: moment {k}"
    else:
        assert np.abs(a - b).max() <= tol, f"{what}: max {np.abs(a - b).max()}"


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("outlet", [O.FACE_PRESSURE, O.FACE_OUTFLOW, O.FACE_CONVECTIVE])
def test_channel_bitexact(Q, outlet):
    dims = (20, 10, 6)
    solid = channel_solid(dims)
    orc, hip = cases.pair(Q, dims, collision="trt", nu=0.05, periodic=(0, 0, 1), solid=solid)
    out_val = {O.FACE_PRESSURE: (1.002,), O.FACE_OUTFLOW: None, O.FACE_CONVECTIVE: (0.05,)}[outlet]
    set_faces(orc, hip, [(O.XNEG, O.FACE_VELOCITY, (0.03, 0.004, -0.002)), (O.XPOS, outlet, out_val)])
    done = 0
    for k in (1, 1, 5, 24):
        orc.step(k)
        hip.run(k)
        done += k
        assert_state(orc, hip, solid, f"D3Q{Q} outlet {KIND[outlet]} step {done}")


@pytest.mark.parametrize("outlet", [O.FACE_PRESSURE, O.FACE_CONVECTIVE])
def test_d3q27_obstacles_next_to_open_faces_bitexact(outlet):
    """D3Q27 fp32 (the odd-step dead-slot fill is active) with solid blocks
    one cell from the inlet and touching the outlet face: solid lanes of
    active tiles under the open faces' wrapped addressing, next to the
    stale-slot capture and restore."""
    dims = (70, 10, 6)
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = solid[:, ny - 1, :] = 1
    solid[1:4, 3:6, 1:4] = 1             # one cell from the inlet face
    solid[2:5, 4:8, nx - 3:nx] = 1       # reaching the outlet face
    solid[:, 2:4, 30:34] = 1             # a pillar mid-channel
    solid = solid.reshape(-1)
    orc, hip = cases.pair(27, dims, collision="trt", nu=0.05, periodic=(0, 0, 1), solid=solid)
    out_val = {O.FACE_PRESSURE: (1.001,), O.FACE_CONVECTIVE: (0.05,)}[outlet]
    set_faces(orc, hip, [(O.XNEG, O.FACE_VELOCITY, (0.03, 0.002, 0.0)), (O.XPOS, outlet, out_val)])
    done = 0
    for k in (1, 1, 5, 12):
        orc.step(k)
        hip.run(k)
        done += k
        assert_state(orc, hip, solid, f"D3Q27 obstacles, outlet {KIND[outlet]} step {done}")


def test_open_edges_and_corners_bitexact():
    """Four open faces meeting at edges (z periodic, no walls): every rule
    at once, in face order, over cells shared by two faces."""
    dims = (14, 12, 4)
    solid = np.zeros(int(np.prod(dims)), np.uint8)
    orc, hip = cases.pair(19, dims, collision="trt", nu=0.04, periodic=(0, 0, 1), solid=solid)
    set_faces(orc, hip, [(O.XNEG, O.FACE_VELOCITY, (0.02, 0.01, 0.0)), (O.XPOS, O.FACE_OUTFLOW, None),
                         (O.YNEG, O.FACE_PRESSURE, (1.001,)), (O.YPOS, O.FACE_CONVECTIVE, (0.1,))])
    done = 0
    for k in (1, 2, 9):
        orc.step(k)
        hip.run(k)
        done += k
        assert_state(orc, hip, solid, f"edges step {done}")


@pytest.mark.parametrize("Q", [19, 27])
def test_forced_zou_he_bitexact(Q):
    """Body force: the Zou-He faces prescribe the raw momentum u - F/(2 rho)."""
    dims = (16, 8, 5)
    solid = channel_solid(dims)
    orc, hip = cases.pair(Q, dims, collision="trt", nu=0.05, periodic=(0, 0, 1), solid=solid,
                          force=(3e-6, -1e-6, 2e-6))
    set_faces(orc, hip, [(O.XNEG, O.FACE_VELOCITY, (0.02, 0.0, 0.003)), (O.XPOS, O.FACE_PRESSURE, (1.0,))])
    for k in (3, 10):
        orc.step(k)
        hip.run(k)
        assert_state(orc, hip, solid, f"forced D3Q{Q}")


def test_bgk_z_faces_bitexact():
    dims = (6, 8, 16)
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, :, 0] = solid[:, :, nx - 1] = 1
    solid = solid.reshape(-1)
    orc, hip = cases.pair(19, dims, collision="bgk", nu=0.03, periodic=(0, 1, 0), solid=solid)
    set_faces(orc, hip, [(O.ZNEG, O.FACE_VELOCITY, (0.0, 0.005, 0.025)), (O.ZPOS, O.FACE_OUTFLOW, None)])
    for k in (2, 7):
        orc.step(k)
        hip.run(k)
        assert_state(orc, hip, solid, "bgk z faces")


def test_central_moment_within_1e6():
    dims = (16, 10, 5)
    solid = channel_solid(dims)
    orc, hip = cases.pair(27, dims, collision="cm", nu=0.02, periodic=(0, 0, 1), solid=solid)
    set_faces(orc, hip, [(O.XNEG, O.FACE_VELOCITY, (0.03, 0.0, 0.0)), (O.XPOS, O.FACE_PRESSURE, (1.0,))])
    orc.step(20)
    hip.run(20)
    assert_state(orc, hip, solid, "cm", exact=False, tol=1e-6)


@pytest.mark.parametrize("storage,tol", [("f16", 2e-4), ("bf16", 2e-3)])
def test_16bit_storage_tracks_f32(storage, tol):
    dims = (20, 10, 4)
    solid = channel_solid(dims)
    runs = []
    for st in ("f32", storage):
        s = L.Solver(19, dims, nu=0.05, storage=st, periodic=(False, False, True), solid=solid)
        s.set_face_bc("x-", "velocity", (0.03, 0.0, 0.0))
        s.set_face_bc("x+", "pressure", (1.0,))
        s.run(30)
        runs.append(s.moments()[1])
        s.close()
    fl = solid == 0
    assert np.abs(runs[0][fl] - runs[1][fl]).max() < tol


def test_captured_span_matches_eager_span():
    """A rotating body in an open channel: the span graph (two steps per
    replay, stash generations included in its key) equals per-step stepping."""
    dims = (24, 16, 16)
    solid = channel_solid(dims)
    th = np.linspace(0, 2 * np.pi, 24, endpoint=False)
    pos = np.stack([12 + 3 * np.cos(th), 8 + 3 * np.sin(th), np.full_like(th, 8.0)], 1)
    w = np.ones(len(th))
    c = np.array([12.0, 8.0, 8.0])
    runs = []
    for graph in (False, True):
        s = L.Solver(19, dims, nu=0.05, periodic=(False, False, True), solid=solid)
        s.set_face_bc("x-", "velocity", (0.02, 0.0, 0.0))
        s.set_face_bc("x+", "convective", (0.05,))
        s.set_rotating_body(pos, w, c, (0.0, 0.0, 0.01))
        tq, _ = s.run_with_ibm(5, graph=graph)
        tq2, _ = s.run_with_ibm(6, graph=graph)
        runs.append((s.populations(), tq, tq2))
        s.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])


def test_refusals():
    dims = (8, 8, 8)
    s = L.Solver(19, dims, nu=0.05, periodic=(True, False, False))
    with pytest.raises(L.LbmError):
        s.set_face_bc("x-", "outflow")
    with pytest.raises(L.LbmError):  # fluid on the y and z faces, none covered
        s.run(1)
    s.set_face_bc("y-", "velocity", (0.0, 0.01, 0.0))
    s.set_face_bc("y+", "outflow")
    with pytest.raises(L.LbmError):  # z faces still uncovered
        s.run(1)
    s.set_face_bc("z-", "outflow")
    s.set_face_bc("z+", "outflow")
    s.run(2)
    s.enable_scalar_distribution(0.01)
    with pytest.raises(L.LbmError):
        s.run(1)
    s.close()


def duct_profile(n1, n2, umax=0.03):
    c1 = np.arange(n1) + 0.5
    c2 = np.arange(n2) + 0.5
    v = np.zeros((n2, n1, 3))
    v[:, :, 0] = umax * np.outer(4 * c2 * (n2 - c2) / n2**2, 4 * c1 * (n1 - c1) / n1**2)
    v[:, :, 2] = 0.002 * np.outer(np.ones(n2), np.linspace(-1, 1, n1))
    return v.reshape(-1, 3).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("Q", [19, 27])
def test_inlet_profile_bitexact(Q):
    """set_inlet_profile (solver.rs:3322-3350): a duct profile on the inlet,
    with a patch on the same face taking its own rule over its rect."""
    dims = (20, 10, 6)
    nx, ny, nz = dims
    solid = channel_solid(dims)
    orc, hip = cases.pair(Q, dims, collision="trt", nu=0.05, periodic=(0, 0, 1), solid=solid)
    set_faces(orc, hip, [(O.XNEG, O.FACE_VELOCITY, (0.01, 0.0, 0.0)), (O.XPOS, O.FACE_PRESSURE, (1.0,))])
    prof = duct_profile(ny, nz)
    orc.set_inlet_profile(O.XNEG, prof)
    hip.set_inlet_profile("x-", prof)
    patch = [(O.XNEG, (3, 1), (4, 3), O.FACE_VELOCITY, (0.02, 0.0, 0.0))]
    orc.set_face_patches(patch)
    hip.set_face_patches([("x-", (3, 1), (4, 3), "velocity", (0.02, 0.0, 0.0))])
    done = 0
    for k in (1, 1, 6, 11):
        orc.step(k)
        hip.run(k)
        done += k
        assert_state(orc, hip, solid, f"D3Q{Q} inlet profile step {done}")
    hip.set_inlet_profile("x-", None)
    orc.set_inlet_profile(O.XNEG, None)
    orc.step(3)
    hip.run(3)
    assert_state(orc, hip, solid, f"D3Q{Q} profile cleared")


def test_inlet_profile_refusals():
    dims = (10, 6, 4)
    s = L.Solver(19, dims, nu=0.05, periodic=(False, True, True))
    s.set_face_bc("x-", "pressure", (1.0,))
    with pytest.raises(L.LbmError):  # not a velocity face
        s.set_inlet_profile("x-", np.zeros((24, 3)))
    s.set_face_bc("x-", "velocity", (0.01, 0.0, 0.0))
    with pytest.raises(ValueError):  # wrong size
        s.set_inlet_profile("x-", np.zeros((23, 3)))
    bad = np.zeros((24, 3))
    bad[5, 0] = np.nan
    with pytest.raises(L.LbmError):
        s.set_inlet_profile("x-", bad)
    s.set_inlet_profile_with("x-", lambda c1, c2: (0.01 * (c1 + 1) / 6, 0.0, 0.0))
    s.set_face_bc("x+", "outflow")
    s.run(4)
    s.close()
