"""3D lid-driven cavity outputs: the centreline velocity profiles of the
reference's cavity test (crates/lbm-core/tests/t15_5_cavity3d.rs:183-243):
u/U along the vertical centreline (x = y = 1/2, z varying) and w/U along the
horizontal one (y = z = 1/2, x varying), each sampled at the 17 fractional
positions of that test's tables, trilinearly between cell centres at
physical position 0.5 + frac (n - 2) (the walls are the one-cell rims)."""
from __future__ import annotations

import numpy as np

# the fractional positions of the reference's two 17-point profile tables
# (t15_5_cavity3d.rs:37-75): vertical line z, horizontal line x
Z_FRACS = (0.0, 0.0547, 0.0625, 0.0703, 0.1016, 0.1719, 0.2813, 0.4531, 0.5, 0.6172, 0.7344, 0.8516, 0.9531,
           0.9609, 0.9688, 0.9766, 1.0)
X_FRACS = (0.0, 0.0625, 0.0703, 0.0781, 0.0938, 0.1563, 0.2266, 0.2344, 0.5, 0.8047, 0.8594, 0.9063, 0.9453,
           0.9531, 0.9609, 0.9688, 1.0)


def _sample(field, n, x, y, z):
    """Trilinear sample of a compact (z, y, x) field at a physical position,
    the corner cells clamped to the fluid interior 1..n-2
    (t15_5_cavity3d.rs:183-214)."""
    f = np.asarray(field, np.float64).reshape(-1)
    x0 = int(min(max(np.floor(x), 1.0), n - 2))
    y0 = int(min(max(np.floor(y), 1.0), n - 2))
    z0 = int(min(max(np.floor(z), 1.0), n - 2))
    x1, y1, z1 = min(x0 + 1, n - 2), min(y0 + 1, n - 2), min(z0 + 1, n - 2)
    tx, ty, tz = x - x0, y - y0, z - z0

    def at(i, j, k):
        return f[(k * n + j) * n + i]

    c00 = (1.0 - tx) * at(x0, y0, z0) + tx * at(x1, y0, z0)
    c10 = (1.0 - tx) * at(x0, y1, z0) + tx * at(x1, y1, z0)
    c01 = (1.0 - tx) * at(x0, y0, z1) + tx * at(x1, y0, z1)
    c11 = (1.0 - tx) * at(x0, y1, z1) + tx * at(x1, y1, z1)
    c0 = (1.0 - ty) * c00 + ty * c10
    c1 = (1.0 - ty) * c01 + ty * c11
    return (1.0 - tz) * c0 + tz * c1


def centrelines(ux, uz, n: int, u_lid: float):
    """(z_fracs, u/U, x_fracs, w/U) of an n^3 cavity's velocity fields; the
    end points carry the wall values (0 at the bottom, 1 under the lid, 0 at
    the side walls) as the reference's line samplers do
    (t15_5_cavity3d.rs:220-243)."""
    pos = lambda frac: 0.5 + frac * (n - 2)  # noqa: E731
    u = []
    for zf in Z_FRACS:
        if zf <= 0.0:
            u.append(0.0)
        elif zf >= 1.0:
            u.append(1.0)
        else:
            u.append(_sample(ux, n, pos(0.5), pos(0.5), pos(zf)) / u_lid)
    w = []
    for xf in X_FRACS:
        if xf <= 0.0 or xf >= 1.0:
            w.append(0.0)
        else:
            w.append(_sample(uz, n, pos(xf), pos(0.5), pos(0.5)) / u_lid)
    return np.array(Z_FRACS), np.array(u), np.array(X_FRACS), np.array(w)
