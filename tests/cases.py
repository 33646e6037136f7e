"""Shared scenario builders for the parity tests: the same seeded inputs go
to the oracle (checker) and to the HIP library (product)."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

COLL = {"bgk": O.BGK, "trt": O.TRT, "cm": O.CM}


def splitmix64(seed):
    """SplitMix64 stream (Steele, Lea, Flood 2014) for seeded inputs."""
    state = np.uint64(seed)
    mask = np.uint64(0xFFFFFFFFFFFFFFFF)
    while True:
        with np.errstate(over="ignore"):
            state = (state + np.uint64(0x9E3779B97F4A7C15)) & mask
            z = state
            z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & mask
            z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & mask
            z = z ^ (z >> np.uint64(31))
        yield float(z >> np.uint64(11)) / float(1 << 53)


def seeded_uniform(seed, n):
    g = splitmix64(seed)
    return np.array([next(g) for _ in range(n)])


def coords(dims):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x.reshape(-1), y.reshape(-1), z.reshape(-1)


def cavity_masks(n, u_lid=0.05):
    """3D lid-driven cavity: six wall rims, ZPos lid moving +x
    (t15_5_cavity3d.rs:148-181)."""
    u = np.zeros((6, 3))
    u[5] = [u_lid, 0.0, 0.0]
    return O.build_wall_rims((n, n, n), [1] * 6, u)


def tgv_field(dims, u0=0.02, seed=None):
    nx, ny, nz = dims
    x, y, z = coords(dims)
    kx, ky, kz = 2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz
    u = np.stack([u0 * np.sin(kx * x) * np.cos(ky * y) * np.cos(kz * z),
                  -u0 * np.cos(kx * x) * np.sin(ky * y) * np.cos(kz * z),
                  0.0 * x], 1)
    rho = 1.0 + 0.001 * np.cos(kx * x) * np.cos(ky * y)
    return rho.astype(np.float32).astype(np.float64), u.astype(np.float32).astype(np.float64)


def pair(Q, dims, *, collision="trt", nu=0.05, periodic=(0, 0, 0), solid=None, wall_u=None, force=(0, 0, 0),
         storage="f32", omega_shear=None):
    """Build (oracle f32 solver, HIP solver) on identical inputs."""
    import lbmflow_amd as L

    f32 = np.array(force, np.float32).astype(np.float64)
    wu32 = None if wall_u is None else np.asarray(wall_u, np.float32).astype(np.float64)
    kw = {}
    if omega_shear is not None:
        kw["omega_shear"] = omega_shear
    orc = O.OracleSolver(Q, dims, prec=4, periodic=periodic, collision=COLL[collision], nu=nu, force=f32,
                         solid=solid, wall_u=wu32, **kw)
    hip = L.Solver(Q, dims, nu=nu, collision=collision, storage=storage, periodic=[bool(p) for p in periodic],
                   force=f32, solid=solid, wall_u=wu32, **({"omega_shear": omega_shear} if omega_shear else {}))
    return orc, hip


def tank_geometry(n, *, T=None, H=None, baffle_width=None, baffle_thickness=None):
    """Stirred tank in lattice units (dx = 1): cylinder T, liquid height H,
    four wall baffles (geometry.rs:657-723)."""
    T = float(n - 2) if T is None else float(T)
    H = float(n - 2) if H is None else float(H)
    bw = T / 10.0 if baffle_width is None else baffle_width
    bt = T / 48.0 if baffle_thickness is None else baffle_thickness
    solid = O.tank_mask((n, n, n), 1.0, T, H)
    solid, baffle = O.baffle_mask((n, n, n), 1.0, T, H, bw, bt, solid)
    return solid, (T, H, bw, bt)
