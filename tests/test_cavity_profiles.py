"""Centreline sampling of the cavity outputs (t15_5_cavity3d.rs:183-243),
pinned on fields where trilinear interpolation is exact."""
import numpy as np

from lbmflow_amd import cavity


def _field(n, fn):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return fn(x.reshape(-1).astype(float), y.reshape(-1).astype(float), z.reshape(-1).astype(float))


def test_linear_fields_are_sampled_exactly():
    n, u_lid = 20, 0.05
    ux = _field(n, lambda x, y, z: 0.001 * x - 0.002 * y + 0.003 * z)
    uz = _field(n, lambda x, y, z: -0.004 * x + 0.001 * y + 0.002 * z)
    zf, u, xf, w = cavity.centrelines(ux, uz, n, u_lid)
    p = lambda f: 0.5 + f * (n - 2)  # noqa: E731
    for k in range(1, 16):
        pz, px = p(zf[k]), p(xf[k])
        # inside the clamped interior [1, n - 2] the sample is the linear value
        if 1.0 <= pz <= n - 2:
            assert abs(u[k] - (0.001 * p(0.5) - 0.002 * p(0.5) + 0.003 * pz) / u_lid) < 1e-12
        if 1.0 <= px <= n - 2:
            assert abs(w[k] - (-0.004 * px + 0.001 * p(0.5) + 0.002 * p(0.5)) / u_lid) < 1e-12
    assert u[0] == 0.0 and u[-1] == 1.0 and w[0] == 0.0 and w[-1] == 0.0


def test_positions_below_the_interior_clamp_to_the_first_fluid_cell():
    n = 20
    ux = _field(n, lambda x, y, z: z)
    zf, u, _, _ = cavity.centrelines(ux, ux, n, 1.0)
    # frac 0.0547 -> physical 1.48: floor 1, weight 0.48 between z = 1 and 2
    assert abs(u[1] - (0.5 + 0.0547 * (n - 2))) < 1e-12
