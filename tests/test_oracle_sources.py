"""Pins of the oracle's volume sources (apply_volume_sources_impl,
backend.rs:799-856): the added populations sum to q_cell per cell and carry
q_cell u of momentum (moments of w_q (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u)),
so a closed periodic box gains exactly the mass and momentum the sources
inject; solid cells of the box are skipped."""
import numpy as np

from oracle import oracle as O


def test_mass_and_momentum_injection_is_exact():
    for Q in (19, 27):
        o = O.OracleSolver(Q, (8, 8, 8), prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=0.1)
        o.set_volume_sources([((2, 2, 2), (4, 3, 3), 1e-3), ((5, 5, 5), (5, 6, 6), -4e-4, (0.05, 0.0, -0.02))])
        m0 = o.mass_deviation()
        o.step(10)
        assert abs(o.mass_deviation() - m0 - 10 * (1e-3 - 4e-4)) < 1e-13
        assert abs(o.reduce(2, 0) - 10 * (-4e-4 * 0.05)) < 1e-15
        assert abs(o.reduce(2, 1)) < 1e-15
        assert abs(o.reduce(2, 2) - 10 * (-4e-4 * -0.02)) < 1e-15


def test_solid_cells_of_the_box_are_skipped():
    dims = (8, 8, 8)
    solid = np.zeros(dims[::-1], np.uint8)
    solid[3, 3, 3] = 1  # one of the 2x2x2 box cells
    o = O.OracleSolver(19, dims, prec=8, periodic=(1, 1, 1), collision=O.TRT, nu=0.1, solid=solid.reshape(-1))
    o.set_volume_sources([((2, 2, 2), (3, 3, 3), 8e-4)])
    m0 = o.mass_deviation()
    o.step(5)
    assert abs(o.mass_deviation() - m0 - 5 * 7 * 1e-4) < 1e-13
