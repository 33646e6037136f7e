"""Pins of the oracle's force probe (stream_row kernels.rs:605-615, row
fold backend.rs:521-544): in a body-force-driven periodic channel at steady
state the walls take exactly the momentum the force puts in, so probing
every solid gives F x (fluid cells); an obstacle symmetric in the spanwise
direction feels no spanwise force; reversing a creeping-flow force
reverses the drag."""
import numpy as np

from oracle import oracle as O


def duct(F, probe_all, steps=3000, Q=19):
    dims = (24, 12, 12)
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    cube = np.zeros_like(solid)
    cube[5:8, 5:8, 10:13] = 1
    solid |= cube
    o = O.OracleSolver(Q, dims, prec=8, periodic=(1, 0, 1), collision=O.TRT, nu=0.1, solid=solid.reshape(-1),
                       force=(F, 0.0, 0.0))
    o.set_force_probe((solid if probe_all else cube).reshape(-1))
    o.step(steps)
    return o.probed_force(), int((solid == 0).sum())


def test_probing_every_solid_balances_the_body_force():
    F = 1e-6
    pf, nfluid = duct(F, True)
    assert abs(pf[0] / (F * nfluid) - 1.0) < 1e-4
    assert abs(pf[1]) < 1e-8 * abs(pf[0]) * 1e3 and abs(pf[2]) < 1e-10


def test_creeping_drag_is_odd_in_the_force_and_has_no_spanwise_force():
    """At this Reynolds number (u ~ 1e-5) the drag is linear in F to the
    inertial correction; the obstacle is symmetric in the periodic z
    direction, so the z force vanishes."""
    a, _ = duct(1e-6, False, steps=1500)
    b, _ = duct(-1e-6, False, steps=1500)
    assert a[0] > 0
    assert abs(b[0] + a[0]) < 1e-3 * a[0]
    assert abs(a[2]) < 1e-9 * a[0]


def test_no_probe_reads_zero():
    o = O.OracleSolver(19, (6, 6, 6), prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=0.1)
    o.step(2)
    assert np.all(o.probed_force() == 0.0)
