"""GPU parity of the volume sources (apply_volume_sources_impl,
backend.rs:799-856) against the oracle: fp32 populations bit-exact after odd
and even steps, mass flow and jet, both lattices, with open faces."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

SOURCES = [((2, 2, 2), (4, 3, 3), 1e-3), ((5, 5, 5), (5, 6, 6), -4e-4, (0.05, 0.0, -0.02))]


def f32(src):
    out = []
    for s in SOURCES if src is None else src:
        q = float(np.float32(s[2]))
        if len(s) > 3:
            out.append((s[0], s[1], q, tuple(float(np.float32(v)) for v in s[3])))
        else:
            out.append((s[0], s[1], q))
    return out


def check(orc, hip, solid, what):
    fl = solid == 0
    a = hip.populations()[:, fl].astype(np.float64)
    b = orc.get_f()[:, fl]
    assert np.array_equal(a, b), f"{what}: max {np.abs(a - b).max()}"


@pytest.mark.parametrize("Q", [19, 27])
def test_sources_in_periodic_box_bitexact(Q):
    dims = (10, 9, 8)
    solid = np.zeros(int(np.prod(dims)), np.uint8)
    orc, hip = cases.pair(Q, dims, collision="trt", nu=0.05, periodic=(1, 1, 1), solid=solid)
    src = f32(None)
    orc.set_volume_sources(src)
    hip.set_volume_sources(src)
    for k in (1, 1, 4):
        orc.step(k)
        hip.run(k)
        check(orc, hip, solid, f"D3Q{Q}")
    assert abs(hip.mass_deviation() - orc.mass_deviation()) < 1e-6


def test_sources_with_open_faces_bitexact():
    dims = (20, 10, 6)
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = solid[:, ny - 1, :] = 1
    solid = solid.reshape(-1)
    orc, hip = cases.pair(19, dims, collision="trt", nu=0.05, periodic=(0, 0, 1), solid=solid)
    src = f32([((8, 3, 1), (9, 5, 3), 2e-3, (0.0, 0.04, 0.0))])
    orc.set_volume_sources(src)
    hip.set_volume_sources(src)
    orc.set_face(O.XNEG, O.FACE_VELOCITY, (0.02, 0.0, 0.0))
    orc.set_face(O.XPOS, O.FACE_PRESSURE, (1.0,))
    hip.set_face_bc("x-", "velocity", (0.02, 0.0, 0.0))
    hip.set_face_bc("x+", "pressure", (1.0,))
    for k in (3, 8):
        orc.step(k)
        hip.run(k)
        check(orc, hip, solid, "open + jet")
    hip.set_volume_sources([])
