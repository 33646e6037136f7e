"""GPU parity of the force probe (stream_row kernels.rs:605-615, row fold
backend.rs:521-544): the probed force of every step equals the oracle's
bit for bit (fp32, probed walls at rest), with periodic wrap, with open
faces (accumulated before the face rules), after odd and even steps."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def duct_masks(dims):
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    cube = np.zeros_like(solid)
    cube[5:8, 4:7, 9:12] = 1
    solid |= cube
    return solid.reshape(-1), cube.reshape(-1)


@pytest.mark.parametrize("Q", [19, 27])
def test_probe_in_periodic_duct_bitexact(Q):
    dims = (24, 12, 12)
    solid, cube = duct_masks(dims)
    orc, hip = cases.pair(Q, dims, collision="trt", nu=0.05, periodic=(1, 0, 1), solid=solid, force=(2e-6, 0, 0))
    for who in (orc, hip):
        who.set_force_probe(cube)
    for k in range(6):
        orc.step(1 + k % 2)
        hip.run(1 + k % 2)
        a, b = hip.probed_force(), orc.probed_force()
        assert np.array_equal(a, b), f"step {k}: {a} vs {b}"
    assert abs(a[0]) > 0
    # every solid probed: walls wrap across the periodic x and z axes
    for who in (orc, hip):
        who.set_force_probe(solid)
    orc.step(3)
    hip.run(3)
    assert np.array_equal(hip.probed_force(), orc.probed_force())


def test_probe_with_open_faces_bitexact():
    dims = (28, 12, 6)
    solid, cube = duct_masks(dims)
    orc, hip = cases.pair(19, dims, collision="trt", nu=0.05, periodic=(0, 0, 1), solid=solid)
    for who in (orc, hip):
        who.set_force_probe(cube)
    orc.set_face(O.XNEG, O.FACE_VELOCITY, (0.03, 0.0, 0.0))
    orc.set_face(O.XPOS, O.FACE_OUTFLOW)
    hip.set_face_bc("x-", "velocity", (0.03, 0.0, 0.0))
    hip.set_face_bc("x+", "outflow")
    for k in range(5):
        orc.step(3)
        hip.run(3)
        assert np.array_equal(hip.probed_force(), orc.probed_force())
    assert hip.probed_force()[0] > 0
    hip.set_force_probe(None)
    hip.run(1)
    assert np.all(hip.probed_force() == 0.0)
