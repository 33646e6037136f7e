"""ctypes loader for the CPU oracle (``liblbm_oracle.so``).

TEST INFRASTRUCTURE ONLY.  The oracle restates lbm-core's ``CpuScalar``
lattice-Boltzmann path (TakuTsuzuki/lbmflow) and is used as the checker by
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``.  Nothing in the ``lbmflow_amd`` package imports this module.

Array conventions: compact global order ``cell = (z * ny + y) * nx + x``;
populations are q-major planes ``f[q, cell]`` in deviation form ``f - w``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbm_oracle.so")
_lib = None

BGK, TRT, CM = 0, 1, 2
MAGIC_STD = 3.0 / 16.0
WALE_CW = 0.325  # les.rs:16
# Face::ALL order (lattice.rs:38-45) and FaceBC kinds (params.rs:232-252)
XNEG, XPOS, YNEG, YPOS, ZNEG, ZPOS = range(6)
FACE_CLOSED, FACE_VELOCITY, FACE_PRESSURE, FACE_OUTFLOW, FACE_CONVECTIVE = range(5)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        d = C.POINTER(C.c_double)
        i = C.POINTER(C.c_int)
        u8 = C.POINTER(C.c_uint8)
        vp = C.c_void_p
        L.orc_lattice.argtypes = [C.c_int, C.POINTER(C.c_int8), d, i]
        L.orc_omegas.argtypes = [C.c_int, C.c_double, C.c_double, d, d]
        L.orc_cm_transform.argtypes = [C.c_int, d, d, u8]
        L.orc_new.restype = vp
        L.orc_new.argtypes = [C.c_int, C.c_int, i, i, C.c_int, C.c_double, C.c_double, C.c_double, d, u8, d]
        L.orc_free.argtypes = [vp]
        L.orc_flush_subnormals.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        L.orc_set_cm_factorised.argtypes = [C.c_int]
        L.orc_write_checkpoint.argtypes = [vp, C.c_char_p]
        L.orc_step.argtypes = [vp, C.c_int]
        L.orc_collide_only.argtypes = [vp]
        L.orc_update_moments.argtypes = [vp]
        L.orc_get_f.argtypes = [vp, d]
        L.orc_set_f.argtypes = [vp, d]
        L.orc_get_f32.argtypes = [vp, C.POINTER(C.c_float)]
        L.orc_set_f32.argtypes = [vp, C.POINTER(C.c_float)]
        L.orc_get_moments.argtypes = [vp, d, d, d, d]
        L.orc_set_moments.argtypes = [vp, d, d, d, d]
        L.orc_reduce.restype = C.c_double
        L.orc_reduce.argtypes = [vp, C.c_int, C.c_int]
        L.orc_set_force_field.argtypes = [vp, d]
        L.orc_init_with.argtypes = [vp, d, d]
        L.orc_equilibrium.argtypes = [vp, C.c_double, d, d]
        L.orc_clear_force_field.argtypes = [vp]
        L.orc_get_force_field.argtypes = [vp, d]
        L.orc_set_omega_field.argtypes = [vp, d]
        L.orc_strain.argtypes = [vp, d]
        L.orc_set_face.argtypes = [vp, C.c_int, C.c_int, d]
        L.orc_set_patches.argtypes = [vp, C.c_int, d]
        L.orc_set_inlet_profile.argtypes = [vp, C.c_int, d]
        L.orc_set_probe.argtypes = [vp, u8]
        L.orc_set_sources.argtypes = [vp, C.c_int, d]
        L.orc_probed_force.argtypes = [vp, d]
        L.orc_vgrad.argtypes = [vp, d]
        L.orc_wale.restype = C.c_long
        L.orc_wale.argtypes = [vp, C.c_double, C.c_double, d, d, d]
        L.orc_apply_ibm.argtypes = [vp, C.c_int, d, d, d, d, C.c_int, C.c_double, C.c_int, C.c_double, d, C.c_int]
        L.orc_scalar_enable.argtypes = [vp, C.c_double, C.c_int]
        L.orc_scalar_set.argtypes = [vp, C.c_int, d]
        L.orc_scalar_get.argtypes = [vp, C.c_int, d]
        L.orc_scalar_get_h.argtypes = [vp, C.c_int, d]
        L.orc_build_wall_rims.argtypes = [i, i, d, u8, d]
        L.orc_tank_mask.argtypes = [i, C.c_double, C.c_double, C.c_double, u8]
        L.orc_baffle_mask.argtypes = [i, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, u8, u8]
        L.orc_radial_impeller.argtypes = [i, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_double, C.c_double, C.c_double, C.c_int, d, d, d, d]
        L.orc_stamp_wall_velocity.argtypes = [i, C.c_int, d, d, d]
        L.orc_impeller_surface.argtypes = [i, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                           C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, d, d, d, d]
        L.orc_marker_stencil.argtypes = [d, i, C.c_int, C.POINTER(C.c_long), d]
        L.orc_power_qois.argtypes = [C.c_double] * 5 + [d]
        L.orc_scalar_cv.argtypes = [C.c_long, d, u8, d]
        L.orc_mixing_time.argtypes = [C.c_long, d, d, d]
        L.orc_compartment_cv.argtypes = [i, d, u8, C.c_int, C.c_double, d, C.POINTER(C.c_long)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def set_cm_factorised(on=True):
    """D3Q27 central moments through the factorised 1-D line transforms
    (default) or the reference's generic matrix path (kernels.rs:227-392)."""
    lib().orc_set_cm_factorised(int(bool(on)))


def max_threads():
    """Threads of the oracle's OpenMP regions (omp_get_max_threads)."""
    return int(lib().orc_max_threads())


def flush_subnormals(on=True):
    """x86 FTZ|DAZ on the oracle's threads (long f64 runs; see the C++)."""
    lib().orc_flush_subnormals(int(bool(on)))


def lattice(Q):
    c = np.zeros((Q, 3), np.int8)
    w = np.zeros(Q)
    opp = np.zeros(Q, np.int32)
    npairs = lib().orc_lattice(Q, c.ctypes.data_as(C.POINTER(C.c_int8)), _d(w), _i(opp))
    return c, w, opp, npairs


def omegas(collision, nu, magic=MAGIC_STD):
    op = C.c_double()
    om = C.c_double()
    lib().orc_omegas(collision, nu, magic, C.byref(op), C.byref(om))
    return op.value, om.value


def cm_transform(Q):
    R = np.zeros((Q, Q))
    P = np.zeros((Q, Q))
    basis = np.zeros((Q, 3), np.uint8)
    lib().orc_cm_transform(Q, _d(R), _d(P), _u8(basis))
    return R, P, basis


class OracleSolver:
    """The reference ``Solver`` over ``CpuScalar`` on one monolithic part."""

    def __init__(self, Q, dims, *, prec=4, periodic=(0, 0, 0), collision=TRT, nu=0.1, magic=MAGIC_STD,
                 omega_shear=None, force=(0.0, 0.0, 0.0), solid=None, wall_u=None):
        self.Q = Q
        self.dims = tuple(int(v) for v in dims)
        self.n = self.dims[0] * self.dims[1] * self.dims[2]
        self.prec = prec
        d = np.array(self.dims, np.int32)
        per = np.array([int(bool(p)) for p in periodic], np.int32)
        frc = np.array(force, np.float64)
        if omega_shear is None:
            omega_shear = omegas(collision, nu, magic)[0]
        sol = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(-1)
        wal = None if wall_u is None else np.ascontiguousarray(wall_u, np.float64).reshape(-1)
        self._keep = (sol, wal)
        self.h = lib().orc_new(prec, Q, _i(d), _i(per), collision, nu, magic, omega_shear, _d(frc),
                               _u8(sol) if sol is not None else None, _d(wal) if wal is not None else None)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def collide_only(self):
        lib().orc_collide_only(self.h)

    def update_moments(self):
        lib().orc_update_moments(self.h)

    def get_f(self, dtype=np.float64):
        """Populations f[q][cell]; dtype float32 skips the f64 copy (an f32
        oracle holds f32 values, so nothing is lost)."""
        if np.dtype(dtype) == np.float32:
            out = np.empty((self.Q, self.n), np.float32)
            lib().orc_get_f32(self.h, out.ctypes.data_as(C.POINTER(C.c_float)))
            return out
        out = np.zeros((self.Q, self.n))
        lib().orc_get_f(self.h, _d(out))
        return out

    def set_f(self, f):
        if isinstance(f, np.ndarray) and f.dtype == np.float32:
            f = np.ascontiguousarray(f)
            lib().orc_set_f32(self.h, f.ctypes.data_as(C.POINTER(C.c_float)))
            return
        f = np.ascontiguousarray(f, np.float64)
        lib().orc_set_f(self.h, _d(f))

    def moments(self):
        r, a, b, c = (np.zeros(self.n) for _ in range(4))
        lib().orc_get_moments(self.h, _d(r), _d(a), _d(b), _d(c))
        return r, a, b, c

    def set_moments(self, rho, ux, uy, uz):
        arrs = [np.ascontiguousarray(v, np.float64) for v in (rho, ux, uy, uz)]
        lib().orc_set_moments(self.h, *[_d(v) for v in arrs])

    def reduce(self, kind, axis=0):
        return lib().orc_reduce(self.h, kind, axis)

    def fluid_cells(self):
        return self.reduce(0)

    def mass_deviation(self):
        return self.reduce(1)

    def init_with(self, rho, u):
        rho = np.ascontiguousarray(np.broadcast_to(rho, (self.n,)), np.float64)
        u = np.ascontiguousarray(np.broadcast_to(u, (self.n, 3)), np.float64)
        lib().orc_init_with(self.h, _d(rho), _d(u))

    def equilibrium(self, rho, u):
        out = np.zeros(self.Q)
        uu = np.ascontiguousarray(u, np.float64)
        lib().orc_equilibrium(self.h, float(rho), _d(uu), _d(out))
        return out

    def set_force_field(self, ff):
        ff = np.ascontiguousarray(ff, np.float64).reshape(-1)
        lib().orc_set_force_field(self.h, _d(ff))

    def clear_force_field(self):
        lib().orc_clear_force_field(self.h)

    def force_field(self):
        out = np.zeros(self.n * 3)
        has = lib().orc_get_force_field(self.h, _d(out))
        return out.reshape(self.n, 3) if has else None

    def set_omega_field(self, om):
        if om is None:
            lib().orc_set_omega_field(self.h, None)
        else:
            om = np.ascontiguousarray(om, np.float64)
            lib().orc_set_omega_field(self.h, _d(om))

    def set_volume_sources(self, sources):
        """``StepParams::sources`` (params.rs:262-283): list of
        (lo, hi, q_lu) mass flows or (lo, hi, q_lu, u) jets."""
        rows = []
        for src in sources:
            lo, hi, q = src[0], src[1], src[2]
            u = src[3] if len(src) > 3 else None
            rows.append(list(lo) + list(hi) + [0.0 if u is None else 1.0, q] + list(u if u is not None else (0, 0, 0)))
        v = np.ascontiguousarray(np.array(rows, np.float64).reshape(-1))
        lib().orc_set_sources(self.h, len(rows), _d(v) if len(rows) else None)

    def set_force_probe(self, mask):
        """``Solver::set_force_probe`` (solver.rs:2741-2760) over a compact mask."""
        if mask is None:
            lib().orc_set_probe(self.h, None)
        else:
            self._probe = np.ascontiguousarray(mask, np.uint8).reshape(-1)
            lib().orc_set_probe(self.h, _u8(self._probe))

    def probed_force(self):
        out = np.zeros(3)
        lib().orc_probed_force(self.h, _d(out))
        return out

    def set_face(self, face, kind, value=(0.0, 0.0, 0.0)):
        """Open-face BC of ``face``: value = u (velocity), (rho,) (pressure),
        (u_conv,) (convective)."""
        v = np.zeros(3)
        vv = np.atleast_1d(np.asarray(value, np.float64))
        v[:vv.size] = vv
        if lib().orc_set_face(self.h, face, kind, _d(v)) != 0:
            raise ValueError("open face on a periodic axis or bad face/kind")

    def set_inlet_profile(self, face, values):
        """Solver::set_inlet_profile (solver.rs:3322-3350): values (face
        cells, 3) in face order (c2 * n1 + c1) on a velocity face; None
        clears."""
        if values is None:
            rc = lib().orc_set_inlet_profile(self.h, face, None)
        else:
            v = np.ascontiguousarray(values, np.float64).reshape(-1)
            rc = lib().orc_set_inlet_profile(self.h, face, _d(v))
        if rc != 0:
            raise ValueError("inlet profile on a face that is not a velocity face")

    def set_face_patches(self, patches):
        """Face patches (FacePatch, params.rs:286-296): a list of
        (face, (lo1, lo2), (hi1, hi2), kind, value), rects inclusive over the
        face's tangents in ascending axis order; replaces the list."""
        rows = np.zeros((max(len(patches), 1), 9))
        for k, (face, lo, hi, kind, value) in enumerate(patches):
            vv = np.atleast_1d(np.asarray(value if value is not None else 0.0, np.float64))
            rows[k, :6] = (face, lo[0], lo[1], hi[0], hi[1], kind)
            rows[k, 6:6 + vv.size] = vv
        if lib().orc_set_patches(self.h, len(patches), _d(rows)) != 0:
            raise ValueError("face patch out of range, overlapping, or on a periodic axis")

    def strain(self):
        """Strain rate per cell, [Sxx, Syy, Szz, Sxy, Sxz, Syz] (solver.rs:5393-5447)."""
        out = np.zeros((self.n, 6))
        lib().orc_strain(self.h, _d(out))
        return out

    def velocity_gradient(self):
        """du_i/dx_j per cell as (N, 3, 3) (solver.rs:5490-5581)."""
        out = np.zeros((self.n, 3, 3))
        lib().orc_vgrad(self.h, _d(out))
        return out

    def wale(self, cw=WALE_CW, tau_eff_max=None):
        """WaleLes::update (les.rs:120-211) without installing the field:
        returns (omega, nu_t, clipped cells, max raw nu_t)."""
        om, nt, mx = np.zeros(self.n), np.zeros(self.n), np.zeros(1)
        k = lib().orc_wale(self.h, cw, 0.0 if tau_eff_max is None else tau_eff_max, _d(om), _d(nt), _d(mx))
        return om, nt, int(k), float(mx[0])

    def write_checkpoint(self, path):
        """The reference's v3 rank file (solver.rs:4536-4630) from this
        solver's own buffers (f, ftmp, moments, solid, force field, scalars)."""
        if lib().orc_write_checkpoint(self.h, str(path).encode()) != 0:
            raise OSError(f"cannot write {path}")

    def apply_ibm(self, pos, weight, center, omega, max_iter=3, tol=1e-2, radius=1, relax=1.0, stencil_dim=3):
        """Solver::apply_rotating_ibm (solver.rs:3016-3282).  stencil_dim=2
        is the reference's 2D marker stencil (z axis a single point of
        weight 1) for one-cell-deep boxes."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
        weight = np.ascontiguousarray(weight, np.float64)
        c = np.ascontiguousarray(center, np.float64)
        o = np.ascontiguousarray(omega, np.float64)
        diag = np.zeros(15)
        rc = lib().orc_apply_ibm(self.h, len(weight), _d(pos), _d(weight), _d(c), _d(o), max_iter, tol, radius,
                                 relax, _d(diag), int(stencil_dim))
        if rc != 0:
            raise ValueError("IBM marker stencil is outside the domain")
        return ibm_diag_dict(diag)

    def scalar_enable(self, D, d3q7=False):
        return lib().orc_scalar_enable(self.h, D, int(d3q7))

    def scalar_set(self, k, values):
        v = np.ascontiguousarray(values, np.float64)
        lib().orc_scalar_set(self.h, k, _d(v))

    def scalar_get(self, k):
        out = np.zeros(self.n)
        lib().orc_scalar_get(self.h, k, _d(out))
        return out

    def scalar_get_h(self, k, q):
        out = np.zeros(q * self.n)
        lib().orc_scalar_get_h(self.h, k, _d(out))
        return out.reshape(q, self.n)


def ibm_diag_dict(d):
    return {
        "iterations": int(d[0]), "slip_max": d[1], "slip_rms": d[2], "slip_max_rel": d[3], "slip_rms_rel": d[4],
        "torque": d[5:8].copy(), "fluid_force": d[8:11].copy(), "marker_force": d[11:14].copy(),
        "momentum_error_rel": d[14],
    }


# ---- geometry ---------------------------------------------------------------

def build_wall_rims(dims, is_wall, u):
    dims = np.array(dims, np.int32)
    n = int(np.prod(dims))
    solid = np.zeros(n, np.uint8)
    wall_u = np.zeros((n, 3))
    iw = np.array(is_wall, np.int32)
    uu = np.ascontiguousarray(u, np.float64).reshape(6, 3)
    lib().orc_build_wall_rims(_i(dims), _i(iw), _d(uu), _u8(solid), _d(wall_u))
    return solid, wall_u


def tank_mask(dims, dx, diameter, height, solid=None):
    dims = np.array(dims, np.int32)
    if solid is None:
        solid = np.zeros(int(np.prod(dims)), np.uint8)
    lib().orc_tank_mask(_i(dims), dx, diameter, height, _u8(solid))
    return solid


def baffle_mask(dims, dx, diameter, height, width, thickness, solid):
    dims = np.array(dims, np.int32)
    baffle = np.zeros(int(np.prod(dims)), np.uint8)
    lib().orc_baffle_mask(_i(dims), dx, diameter, height, width, thickness, _u8(solid), _u8(baffle))
    return solid, baffle


def radial_impeller(dims, dx, vessel_diameter, liquid_height, diameter, clearance, blades=6, blade_angle_deg=0.0,
                    omega_z=0.0, thickness_fraction=0.04, inner_kind=0):
    dims = np.array(dims, np.int32)
    n = lib().orc_radial_impeller(_i(dims), dx, vessel_diameter, liquid_height, diameter, clearance, blades,
                                  blade_angle_deg, omega_z, thickness_fraction, inner_kind, None, None, None, None)
    pos = np.zeros((n, 3))
    w = np.zeros(n)
    vel = np.zeros((n, 3))
    c = np.zeros(3)
    lib().orc_radial_impeller(_i(dims), dx, vessel_diameter, liquid_height, diameter, clearance, blades,
                              blade_angle_deg, omega_z, thickness_fraction, inner_kind, _d(pos), _d(w), _d(vel),
                              _d(c))
    return pos, w, vel, c


def impeller_surface(dims, dx, vessel_diameter, liquid_height, diameter, clearance, blades=6, blade_angle_deg=0.0,
                     omega_z=0.0, thickness_fraction=0.04, area_per_marker=0.1, seed=20260512):
    """Seeded uniform surface sampling of the blade boxes (restated from its
    specification; the reference has no surface sampler)."""
    dims = np.array(dims, np.int32)
    args = (_i(dims), dx, vessel_diameter, liquid_height, diameter, clearance, blades, blade_angle_deg, omega_z,
            thickness_fraction, area_per_marker, int(seed) & 0xFFFFFFFFFFFFFFFF)
    n = lib().orc_impeller_surface(*args, None, None, None, None)
    pos = np.zeros((n, 3))
    w = np.zeros(n)
    vel = np.zeros((n, 3))
    c = np.zeros(3)
    lib().orc_impeller_surface(*args, _d(pos), _d(w), _d(vel), _d(c))
    return pos, w, vel, c


def stamp_wall_velocity(dims, pos, vel, wall_u):
    dims = np.array(dims, np.int32)
    pos = np.ascontiguousarray(pos, np.float64)
    vel = np.ascontiguousarray(vel, np.float64)
    lib().orc_stamp_wall_velocity(_i(dims), len(pos), _d(pos), _d(vel), _d(wall_u))
    return wall_u


def marker_stencil(p, dims, radius=1):
    p = np.ascontiguousarray(p, np.float64)
    dims = np.array(dims, np.int32)
    cells = np.zeros(64 * 3, np.int64)
    w = np.zeros(64)
    n = lib().orc_marker_stencil(_d(p), _i(dims), radius, cells.ctypes.data_as(C.POINTER(C.c_long)), _d(w))
    return cells[: 3 * n].reshape(n, 3), w[:n]


# ---- QOIs -------------------------------------------------------------------

def power_qois(torque, omega, rho, d, volume):
    out = np.zeros(4)
    rc = lib().orc_power_qois(torque, omega, rho, d, volume, _d(out))
    if rc != 0:
        raise ValueError("invalid power QOI inputs")
    return {"power": out[0], "n_hz": out[1], "np": out[2], "p_over_v": out[3]}


def scalar_cv(values, include):
    v = np.ascontiguousarray(values, np.float64)
    inc = np.ascontiguousarray(include, np.uint8)
    cv = C.c_double()
    ok = lib().orc_scalar_cv(len(v), _d(v), _u8(inc), C.byref(cv))
    return cv.value if ok else None


def mixing_time(t, cv):
    t = np.ascontiguousarray(t, np.float64)
    cv = np.ascontiguousarray(cv, np.float64)
    out = np.zeros(3)
    ok = lib().orc_mixing_time(len(t), _d(t), _d(cv), _d(out))
    if not ok:
        return None
    f = lambda v: None if np.isnan(v) else float(v)  # noqa: E731
    return {"cv0": out[0], "t95": f(out[1]), "t99": f(out[2])}


def compartment_cv(dims, values, solid, imp_z=None):
    d = np.array(dims, np.int32)
    v = np.ascontiguousarray(values, np.float64)
    s = np.ascontiguousarray(solid, np.uint8)
    out = np.zeros(3)
    counts = np.zeros(3, np.int64)
    lib().orc_compartment_cv(_i(d), _d(v), _u8(s), int(imp_z is not None), 0.0 if imp_z is None else imp_z,
                             _d(out), counts.ctypes.data_as(C.POINTER(C.c_long)))
    names = ("top", "bulk", "near_impeller")
    return {nm: (None if np.isnan(out[k]) else out[k], int(counts[k])) for k, nm in enumerate(names)}
