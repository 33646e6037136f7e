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


def tank_case(n, *, nu=0.01, tip=0.05, collision="cm", Q=27, sigma=None, markers="lattice", seed=20260512):
    """Small stirred tank (same ratios as the benchmark): geometry, markers
    and a Gaussian tracer near the surface.  markers: "lattice" (the
    reference's volumetric blade lattice) or "surface" (the seeded surface
    sampler, one marker per 0.1 lu^2)."""
    T = 15.0 * n / 16.0
    H = T
    solid, _ = tank_geometry(n, T=T, H=H, baffle_width=T / 10.0, baffle_thickness=T / 48.0)
    D = T / 3.0
    omega = tip / (0.5 * D)
    if markers == "surface":
        pos, w, vel, c = O.impeller_surface((n, n, n), 1.0, T, H, D, T / 3.0, 6, 0.0, omega, 0.04, 0.1, seed)
    else:
        pos, w, vel, c = O.radial_impeller((n, n, n), 1.0, T, H, D, T / 3.0, 6, 0.0, omega)
    x, y, z = coords((n, n, n))
    sig = (8.0 * n / 512.0 * 8.0) if sigma is None else sigma
    cz = 0.8 * H
    blob = np.exp(-((x + 0.5 - n / 2) ** 2 + (y + 0.5 - n / 2) ** 2 + (z + 0.5 - cz) ** 2) / (2 * sig ** 2))
    blob = np.where(solid == 0, blob, 0.0).astype(np.float32).astype(np.float64)
    return {"n": n, "T": T, "H": H, "D": D, "omega": omega, "nu": nu, "solid": solid, "pos": pos, "w": w,
            "c": c, "blob": blob, "collision": collision, "Q": Q}


def run_tank_oracle(tc, steps, prec=4, every=1, scalar_D=None, d3q7=True, cv_every=None, progress=0):
    s = O.OracleSolver(tc["Q"], (tc["n"],) * 3, prec=prec, collision=COLL[tc["collision"]], nu=tc["nu"],
                       solid=tc["solid"])
    k = None
    if scalar_D is not None:
        k = s.scalar_enable(scalar_D, d3q7)
        s.scalar_set(k, tc["blob"])
    torque, cv = [], []
    fl = (tc["solid"] == 0).astype(np.uint8)
    for step in range(1, steps + 1):
        if progress and step % progress == 0:
            print(f"[oracle] step {step}/{steps}", flush=True)
        s.clear_force_field()
        d = s.apply_ibm(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
        s.step()
        if step % every == 0:
            torque.append(-d["torque"][2])
        if k is not None and step % (cv_every or every) == 0:
            cv.append(O.scalar_cv(s.scalar_get(k), fl))
    return np.array(torque), np.array(cv), s


def run_tank_gpu(tc, steps, storage="f32", every=1, scalar_D=None, d3q7=True, cv_every=None):
    import lbmflow_amd as L

    n = tc["n"]
    s = L.Solver(tc["Q"], (n, n, n), nu=tc["nu"], collision=tc["collision"], storage=storage, solid=tc["solid"])
    s.set_rotating_body(tc["pos"], tc["w"], tc["c"], (0.0, 0.0, tc["omega"]))
    k = None
    if scalar_D is not None:
        k = s.enable_scalar_distribution(scalar_D, "d3q7" if d3q7 else "hydro")
        s.set_scalar_concentration(k, tc["blob"])
    torque, cv = [], []
    for step in range(1, steps + 1):
        s.clear_body_force_field()
        d = s.apply_rotating_ibm()
        s.step()
        if step % every == 0:
            torque.append(-d["torque"][2])
        if k is not None and step % (cv_every or every) == 0:
            cv.append(s.scalar_cv(k))
    return np.array(torque), np.array(cv), s


def np_from_torque(torque, tc, steps, every=1):
    """Last-half mean shaft torque (runner.rs:1514-1530) -> Np (qoi.rs:477-505)."""
    steps_rec = np.arange(1, len(torque) + 1) * every
    win = torque[steps_rec >= steps // 2]
    tq = win.mean() if len(win) else torque.mean()
    q = O.power_qois(tq, tc["omega"], 1.0, tc["D"], 1.0)
    return q["np"], tq


def links_from_solid(solid, dims, Q, periodic=(0, 0, 0)):
    c, _, _, _ = O.lattice(Q)
    nx, ny, nz = dims
    s3 = solid.reshape(nz, ny, nx).astype(bool)
    out = []
    for q in range(1, Q):
        # source cell of direction q at (x, y, z) is (x, y, z) - c_q
        src = np.roll(s3, shift=(int(c[q, 2]), int(c[q, 1]), int(c[q, 0])), axis=(0, 1, 2))
        hit = (~s3) & src
        cells = np.flatnonzero(hit.reshape(-1))
        out.extend((int(i), q) for i in cells)
    out.sort()
    return np.array([o[0] for o in out], np.int64), np.array([o[1] for o in out], np.uint8)
