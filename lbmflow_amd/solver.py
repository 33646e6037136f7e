"""Host-side mirror of lbm-core's ``Solver`` over the MI355X backend.

The names, argument meanings and error behaviour follow the reference's
``Solver`` (crates/lbm-core/src/solver.rs) and ``Backend`` trait
(crates/lbm-core/src/backend.rs:126-363) for the single-phase stirred-tank
path: build from solid/wall-velocity masks, step/run, init_with, body-force
fields, rotating IBM, scalar distributions and the f64 reductions.  Every
call goes through the C ABI of ``liblbm_hip.so``; there is no CPU path.

Array conventions: compact global order ``cell = (z*ny + y)*nx + x``,
populations q-major ``f[q, cell]`` in deviation form ``f - w``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as _L

MAGIC_STD = 3.0 / 16.0  # CollisionKind::MAGIC_STD, params.rs:202
WALE_CW = 0.325  # WALE_CW, les.rs:16


def omegas(collision: str, nu: float, magic: float = MAGIC_STD):
    """``CollisionKind::omegas`` (params.rs:205-217)."""
    tau = 3.0 * nu + 0.5
    omega_p = 1.0 / tau
    if collision == "trt":
        lam_p = tau - 0.5
        omega_m = 1.0 / (magic / lam_p + 0.5)
    else:
        omega_m = omega_p
    return omega_p, omega_m


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Context:
    """One device and one HIP stream (calls on it are serialized)."""

    def __init__(self, device: int = 0):
        self.lib = _L.lib()
        h = C.c_void_p()
        rc = self.lib.lbm_ctx_create(device, C.byref(h))
        if rc != 0:
            raise _L.LbmError(rc, f"cannot create a context on device {device}")
        self.h = h

    def check(self, rc):
        if rc != 0:
            msg = self.lib.lbm_last_error(self.h)
            raise _L.LbmError(rc, msg.decode() if msg else "")

    def capabilities(self) -> int:
        v = C.c_uint32()
        self.check(self.lib.lbm_capabilities(self.h, C.byref(v)))
        return v.value

    def stream_copy_gbps(self, nbytes: int = 2 << 30, iterations: int = 20) -> float:
        """Measured copy bandwidth (read + write bytes) on this device."""
        v = C.c_double()
        self.check(self.lib.lbm_stream_copy(self.h, int(nbytes), int(iterations), C.byref(v)))
        return v.value

    def access_copy_gbps(self, width: int, offset_elems: int = 0, nbytes: int = 1 << 30, iterations: int = 5) -> float:
        """Copy with `width`-byte lane accesses (16, 4, 2) starting
        `offset_elems` elements past aligned bases (lbm_access_copy)."""
        v = C.c_double()
        self.check(self.lib.lbm_access_copy(self.h, int(width), int(offset_elems), int(nbytes), int(iterations),
                                            C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class DirectForcingConfig:
    """``DirectForcingConfig`` (rotating_ibm.rs:104-130)."""

    max_iterations: int = 3
    slip_tolerance: float = 1.0e-2
    kernel_radius: int = 1
    relaxation: float = 1.0


def _diag_dict(d):
    return {
        "iterations": d.iterations, "slip_max": d.slip_max, "slip_rms": d.slip_rms,
        "slip_max_rel": d.slip_max_rel, "slip_rms_rel": d.slip_rms_rel, "torque": np.array(d.torque[:]),
        "fluid_force": np.array(d.fluid_force[:]), "marker_force": np.array(d.marker_force[:]),
        "momentum_error_rel": d.momentum_error_rel,
    }


class Solver:
    """A monolithic domain on one GPU (``Solver<L, f32, HipBackend, LocalPeriodic>``)."""

    def __init__(self, lattice: int, dims, *, nu: float, collision: str = "trt", magic: float = MAGIC_STD,
                 omega_shear: float | None = None, storage: str = "f32", periodic=(False, False, False),
                 force=(0.0, 0.0, 0.0), solid=None, wall_u=None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.Q = int(lattice)
        self.dims = tuple(int(v) for v in dims)
        self.n = self.dims[0] * self.dims[1] * self.dims[2]
        self.storage = storage
        self.collision = collision
        self.nu = float(nu)
        self.magic = float(magic)
        self.periodic = tuple(bool(p) for p in periodic)
        self.force = tuple(float(v) for v in force)
        d = np.array(self.dims, np.int32)
        per = np.array([int(bool(p)) for p in periodic], np.int32)
        h = C.c_void_p()
        self.ctx.check(self.lib.lbm_fields_alloc(self.ctx.h, self.Q, _L.STORAGE[storage], _ip(d), _ip(per),
                                                 C.byref(h)))
        self.h = h
        op, om = omegas(collision, self.nu, magic)
        self.params = _L.StepParams()
        self.params.collision = _L.COLLISION[collision]
        self.params.omega_p = op
        self.params.omega_m = om
        self.params.omega_shear = op if omega_shear is None else float(omega_shear)
        for a in range(3):
            self.params.force[a] = float(force[a])
        self.ctx.check(self.lib.lbm_set_params(self.h, C.byref(self.params)))
        if solid is not None or wall_u is not None or not all(periodic):
            self.set_masks(solid, wall_u)
        self._ibm_key = None
        self.scalars = []
        self.sources = []
        self.faces = [("closed", None)] * 6
        self.patches = []

    # ---- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_fields_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        self.ctx.check(rc)

    # ---- masks / geometry -------------------------------------------------
    def set_masks(self, solid=None, wall_u=None):
        s = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(-1)
        w = None if wall_u is None else np.ascontiguousarray(wall_u, np.float32).reshape(-1)
        if s is not None and s.size != self.n:
            raise ValueError("solid mask has the wrong size")
        if w is not None and w.size != 3 * self.n:
            raise ValueError("wall velocity has the wrong size")
        self._chk(self.lib.lbm_set_masks(self.h, None if s is None else _u8p(s), None if w is None else _fp(w)))

    def build_wall_rims(self, is_wall, u):
        """``build_wall_rims`` (solver.rs:1357-1406) on the device."""
        iw = np.array([int(bool(v)) for v in is_wall], np.int32)
        uu = np.ascontiguousarray(u, np.float64).reshape(18)
        self._chk(self.lib.lbm_geom_wall_rims(self.h, _ip(iw), _dp(uu)))

    def build_stirred_tank(self, dx, vessel_diameter, liquid_height, baffle_count=4, baffle_width=0.0,
                           baffle_thickness=0.0):
        """Tank mask + four wall baffles (geometry.rs:657-723) on the device."""
        self._chk(self.lib.lbm_geom_stirred_tank(self.h, dx, vessel_diameter, liquid_height, baffle_count,
                                                 baffle_width, baffle_thickness))

    def solid(self):
        out = np.zeros(self.n, np.uint8)
        self._chk(self.lib.lbm_read_solid(self.h, _u8p(out)))
        return out

    def boundary_links(self):
        cnt = C.c_int64()
        self._chk(self.lib.lbm_boundary_links(self.h, C.byref(cnt), None, None, 0))
        cells = np.zeros(cnt.value, np.int64)
        qs = np.zeros(cnt.value, np.uint8)
        self._chk(self.lib.lbm_boundary_links(self.h, C.byref(cnt), cells.ctypes.data_as(C.POINTER(C.c_int64)),
                                              _u8p(qs), cnt.value))
        return cells, qs

    # ---- state ------------------------------------------------------------
    def set_populations(self, f):
        """``Backend::stage_in`` (q-major compact deviation populations)."""
        f = np.ascontiguousarray(f, np.float32).reshape(-1)
        if f.size != self.Q * self.n:
            raise ValueError("population array has the wrong size")
        self._chk(self.lib.lbm_stage_in(self.h, _fp(f)))

    def populations(self):
        """``Backend::stage_out``; solid cells read as zero."""
        out = np.zeros((self.Q, self.n), np.float32)
        self._chk(self.lib.lbm_stage_out(self.h, _fp(out)))
        return out

    def init_with(self, rho, u):
        """``Solver::init_with`` (solver.rs:2474-2623) from per-cell rho and u."""
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, np.float32), (self.n,)))
        uu = np.ascontiguousarray(np.broadcast_to(np.asarray(u, np.float32), (self.n, 3)))
        self._chk(self.lib.lbm_init_with(self.h, _fp(r), _fp(uu)))

    def set_omega_field(self, omega):
        if omega is None:
            self._chk(self.lib.lbm_set_omega_field(self.h, None))
        else:
            om = np.ascontiguousarray(omega, np.float32).reshape(-1)
            self._chk(self.lib.lbm_set_omega_field(self.h, _fp(om)))

    def set_volume_sources(self, sources):
        """``StepParams::sources`` (params.rs:262-283): a list of
        (lo, hi, q_lu) mass flows and (lo, hi, q_lu, u) jets over inclusive
        cell boxes."""
        arr = (_L.VolumeSource * max(len(sources), 1))()
        for k, src in enumerate(sources):
            e = arr[k]
            for c in range(3):
                e.lo[c] = int(src[0][c])
                e.hi[c] = int(src[1][c])
            e.q_lu = float(src[2])
            e.kind = 1 if len(src) > 3 else 0
            if len(src) > 3:
                for c in range(3):
                    e.u[c] = float(src[3][c])
        self._chk(self.lib.lbm_set_volume_sources(self.h, len(sources), arr))
        self.sources = [tuple(src) for src in sources]

    def set_force_probe(self, mask):
        """``Solver::set_force_probe`` (solver.rs:2741-2760): solids whose
        momentum-exchange force is accumulated each step (None removes)."""
        if mask is None:
            self._chk(self.lib.lbm_set_force_probe(self.h, None))
            return
        m = np.ascontiguousarray(mask, np.uint8).reshape(-1)
        if m.size != self.n:
            raise ValueError("probe mask has the wrong size")
        self._chk(self.lib.lbm_set_force_probe(self.h, _u8p(m)))

    def probed_force(self):
        """``Solver::probed_force`` (solver.rs:5065-5071): the most recent step's."""
        out = np.zeros(3)
        self._chk(self.lib.lbm_probed_force(self.h, _dp(out)))
        return out

    FACES = {"x-": 0, "x+": 1, "y-": 2, "y+": 3, "z-": 4, "z+": 5}
    FACE_KINDS = {"closed": 0, "velocity": 1, "pressure": 2, "outflow": 3, "convective": 4}

    def set_face_bc(self, face, kind: str, value=None):
        """``GlobalSpec::faces[face]`` (params.rs:232-252): face 0..5 or
        "x-".."z+"; kind closed | velocity (value = u) | pressure (value =
        rho) | outflow | convective (value = u_conv)."""
        f = self.FACES[face] if isinstance(face, str) else int(face)
        v = np.zeros(3)
        if value is not None:
            vv = np.atleast_1d(np.asarray(value, np.float64))
            v[:vv.size] = vv
        self._chk(self.lib.lbm_set_face_bc(self.h, f, self.FACE_KINDS[kind], _dp(v)))
        self.faces[f] = (kind, None if kind in ("closed", "outflow") else v.copy())

    def set_inlet_profile(self, face, values):
        """``Solver::set_inlet_profile`` (solver.rs:3322-3350): ``values`` of
        shape (face cells, 3) in face order (c2 * n1 + c1, t1 and t2 the
        face's other axes ascending) on a velocity face; None clears."""
        f = self.FACES[face] if isinstance(face, str) else int(face)
        if values is None:
            self._chk(self.lib.lbm_set_inlet_profile(self.h, f, None))
            return
        v = np.ascontiguousarray(values, np.float64).reshape(-1)
        ext = self.dims
        ax = f // 2
        t1, t2 = (1, 2) if ax == 0 else ((0, 2) if ax == 1 else (0, 1))
        if v.size != 3 * ext[t1] * ext[t2]:
            raise ValueError("profile must cover the whole face")
        self._chk(self.lib.lbm_set_inlet_profile(self.h, f, _dp(v)))

    def set_inlet_profile_with(self, face, profile):
        """``Solver::set_inlet_profile_with`` (solver.rs:3352-3366):
        ``profile(c1, c2)`` at every face cell's tangent coordinates."""
        f = self.FACES[face] if isinstance(face, str) else int(face)
        ax = f // 2
        t1, t2 = (1, 2) if ax == 0 else ((0, 2) if ax == 1 else (0, 1))
        vals = [profile(c1, c2) for c2 in range(self.dims[t2]) for c1 in range(self.dims[t1])]
        self.set_inlet_profile(f, np.asarray(vals, np.float64))

    def set_face_patches(self, patches):
        """``GlobalSpec::face_patches`` (FacePatch, params.rs:286-296): a list
        of (face, (lo1, lo2), (hi1, hi2), kind, value) with inclusive rects
        over the face's two other axes in ascending order, kind and value as
        for :meth:`set_face_bc`; replaces the list."""
        arr = (_L.FacePatch * max(len(patches), 1))()
        for k, (face, lo, hi, kind, value) in enumerate(patches):
            e = arr[k]
            e.face = self.FACES[face] if isinstance(face, str) else int(face)
            e.lo[0], e.lo[1] = int(lo[0]), int(lo[1])
            e.hi[0], e.hi[1] = int(hi[0]), int(hi[1])
            e.kind = self.FACE_KINDS[kind]
            if value is not None:
                vv = np.atleast_1d(np.asarray(value, np.float64))
                for c in range(vv.size):
                    e.value[c] = float(vv[c])
        self._chk(self.lib.lbm_set_face_patches(self.h, len(patches), arr))
        self.patches = [((self.FACES[fc] if isinstance(fc, str) else int(fc)), tuple(lo), tuple(hi), kind, value)
                        for fc, lo, hi, kind, value in patches]

    def wale_update(self, cw: float = WALE_CW, tau_eff_max: float | None = None, want_nu_t: bool = False):
        """One ``WaleLes::update`` (les.rs:120-211): install the WALE omega
        field of the current state.  Returns the diagnostics dict
        (``WaleLesDiagnostics``, les.rs:18-31) and, with ``want_nu_t``, the
        eddy viscosity per cell (after the clip)."""
        d = _L.WaleDiag()
        nut = np.zeros(self.n, np.float32) if want_nu_t else None
        self._chk(self.lib.lbm_wale_update(self.h, cw, 0.0 if tau_eff_max is None else tau_eff_max,
                                           _fp(nut) if nut is not None else None, C.byref(d)))
        diag = {
            "clipped_cells": int(d.clipped_cells),
            "clipped_fraction": d.clipped_fraction,
            "max_nu_t_before_clipping": d.max_nu_t_before_clipping,
            "tau_eff_max": None if tau_eff_max is None else d.tau_eff_max,
            "nu_t_max": None if tau_eff_max is None else d.nu_t_max,
        }
        return (diag, nut) if want_nu_t else diag

    def set_wale(self, enabled: bool, cw: float = WALE_CW, tau_eff_max: float | None = None):
        """Recompute the WALE omega field on the device before every step
        (``GpuSolver::set_wale``, gpu/solver.rs:60-73)."""
        self._chk(self.lib.lbm_set_wale(self.h, int(bool(enabled)), cw,
                                        0.0 if tau_eff_max is None else tau_eff_max))

    def omega_field(self):
        """The installed per-cell omega_plus field, or None
        (``GpuSolver::gather_wale_omega``, gpu/solver.rs:228-235)."""
        out = np.zeros(self.n, np.float32)
        present = C.c_int()
        self._chk(self.lib.lbm_gather_omega_field(self.h, _fp(out), C.byref(present)))
        return out if present.value else None

    def set_body_force_field(self, ff):
        """``Solver::set_body_force_field`` (solver.rs:2776-2801)."""
        ff = np.ascontiguousarray(ff, np.float32).reshape(-1)
        if ff.size != 3 * self.n:
            raise ValueError("force field has the wrong size")
        self._chk(self.lib.lbm_set_force_field(self.h, _fp(ff)))

    def clear_body_force_field(self):
        """``Solver::clear_body_force_field`` (solver.rs:2805-2822)."""
        self._chk(self.lib.lbm_clear_force_field(self.h))

    def body_force_field(self):
        out = np.zeros((self.n, 3), np.float32)
        present = C.c_int()
        self._chk(self.lib.lbm_read_force_field(self.h, _fp(out), C.byref(present)))
        return out if present.value else None

    # ---- stepping -----------------------------------------------------------
    def step(self):
        self._chk(self.lib.lbm_run_span(self.h, 1))

    def run(self, steps: int):
        self._chk(self.lib.lbm_run_span(self.h, int(steps)))

    def synchronize(self):
        self._chk(self.lib.lbm_synchronize(self.h))

    def kernel_timing(self, enable: bool):
        self._chk(self.lib.lbm_kernel_timing(self.h, int(bool(enable))))

    def kernel_timing_read(self):
        ms = C.c_double()
        n = C.c_int64()
        self._chk(self.lib.lbm_kernel_timing_read(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def force_list_size(self) -> int:
        return int(self.lib.lbm_force_list_size(self.h))

    @property
    def time(self) -> int:
        return int(self.lib.lbm_time(self.h))

    # ---- readback / reductions ---------------------------------------------
    def moments(self):
        r, a, b, c = (np.zeros(self.n, np.float32) for _ in range(4))
        self._chk(self.lib.lbm_read_moments(self.h, _fp(r), _fp(a), _fp(b), _fp(c)))
        return r, a, b, c

    def reduce(self, kind: int, axis: int = 0) -> float:
        out = C.c_double()
        self._chk(self.lib.lbm_reduce(self.h, kind, axis, C.byref(out)))
        return out.value

    def fluid_cells(self) -> float:
        return self.reduce(0)

    def mass_deviation(self) -> float:
        return self.reduce(1)

    def total_mass_f64(self) -> float:
        """``Solver::total_mass_f64`` (solver.rs:5114-5135)."""
        return self.fluid_cells() + self.mass_deviation()

    def momentum(self):
        return np.array([self.reduce(2, a) for a in range(3)])

    def device_bytes(self) -> int:
        return int(self.lib.lbm_fields_device_bytes(self.h))

    # ---- rotating IBM ------------------------------------------------------
    def set_rotating_body(self, positions, weights, center, omega, cfg: DirectForcingConfig | None = None):
        cfg = cfg or DirectForcingConfig()
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        w = np.ascontiguousarray(weights, np.float64).reshape(-1)
        c = np.ascontiguousarray(center, np.float64)
        o = np.ascontiguousarray(omega, np.float64)
        ic = _L.IbmConfig(cfg.max_iterations, cfg.slip_tolerance, cfg.kernel_radius, cfg.relaxation)
        self._chk(self.lib.lbm_ibm_set_body(self.h, w.size, _dp(pos), _dp(w), _dp(c), _dp(o), C.byref(ic)))

    def apply_rotating_ibm(self):
        """``Solver::apply_rotating_ibm`` (solver.rs:3016-3282) for the registered body."""
        d = _L.IbmDiag()
        self._chk(self.lib.lbm_ibm_apply(self.h, C.byref(d)))
        return _diag_dict(d)

    def run_with_ibm(self, steps: int, graph: bool = True):
        """``steps`` runner steps (runner.rs:1079-1129: clear the force field,
        rotating IBM, collide-stream, scalars) without host synchronisation
        inside the span.  Returns (torque[steps, 3], diagnostics of the last
        step).  ``graph`` replays pairs of steps from one captured HIP graph."""
        steps = int(steps)
        tq = np.zeros((max(steps, 0), 3))
        d = _L.IbmDiag()
        self._chk(self.lib.lbm_run_ibm_span(self.h, steps, int(bool(graph)), _dp(tq), C.byref(d)))
        return tq, (_diag_dict(d) if steps > 0 else None)

    def autotune_tiles(self, trial_steps: int = 10):
        """Pick the tile visit order of the collide-stream launch per parity
        by timing candidates on a scratch copy of the state
        (lbm_autotune_tiles).  Returns {"zb_even", "zb_odd", "ms_even",
        "ms_odd"}."""
        zb = np.zeros(2, np.uint32)
        ms = np.zeros(2)
        self._chk(self.lib.lbm_autotune_tiles(self.h, int(trial_steps),
                                              zb.ctypes.data_as(C.POINTER(C.c_uint32)), _dp(ms)))
        return {"zb_even": int(zb[0]), "zb_odd": int(zb[1]), "ms_even": float(ms[0]), "ms_odd": float(ms[1])}

    def torque_integral(self, reset: bool = False):
        """Running FP64 sum of the body's torque kept on the device since the
        body was registered or last reset (lbm_ibm_torque_integral): returns
        (sum[3], updates).  The mean shaft torque of a window is
        sum / updates (runner.rs:1514-1530)."""
        out = np.zeros(4)
        self._chk(self.lib.lbm_ibm_torque_integral(self.h, int(bool(reset)), _dp(out)))
        return out[:3].copy(), int(out[3])

    def adjoint_torque(self, steps: int, window: int, snapshots: int = 16, gradients: bool = True):
        """Mean shaft torque J over the last ``window`` of ``steps`` runner
        steps from the current state and, with ``gradients``, its discrete
        adjoint: dJ/dOmega_z of the rotating body and dJ/dnu per cell
        (lbm_adjoint_torque).  The solver ends after the ``steps`` steps."""
        J, dw = C.c_double(), C.c_double()
        fw = C.c_int64()
        dnu = np.zeros(self.n) if gradients else None
        self._chk(self.lib.lbm_adjoint_torque(self.h, int(steps), int(window), int(snapshots), C.byref(J),
                                              C.byref(dw) if gradients else None,
                                              _dp(dnu) if gradients else None, C.byref(fw)))
        out = {"J": J.value, "forward_steps": fw.value}
        if gradients:
            out["dJ_domega"] = dw.value
            out["dJ_dnu"] = dnu
        return out

    # ---- checkpoints (format v3, solver.rs:4536-4830) ------------------------
    def save_checkpoint(self, directory: str) -> dict:
        from . import checkpoint

        return checkpoint.save(self, directory)

    def restore_checkpoint(self, directory: str) -> int:
        from . import checkpoint

        return checkpoint.restore(self, directory)

    # ---- scalar distributions ----------------------------------------------
    def enable_scalar_distribution(self, diffusivity: float, lattice: str = "hydro") -> int:
        idx = C.c_int()
        lat = 7 if lattice == "d3q7" else 0
        self._chk(self.lib.lbm_scalar_enable(self.h, float(diffusivity), lat, C.byref(idx)))
        self.scalars.append(lattice)
        return idx.value

    def set_scalar_concentration(self, index: int, values):
        v = np.ascontiguousarray(values, np.float32).reshape(-1)
        if v.size != self.n:
            raise ValueError(f"scalar concentration length {v.size} does not match cell count {self.n}")
        self._chk(self.lib.lbm_scalar_set(self.h, index, _fp(v)))

    def gather_scalar_concentration(self, index: int):
        out = np.zeros(self.n, np.float32)
        self._chk(self.lib.lbm_scalar_gather(self.h, index, _fp(out)))
        return out

    def scalar_cv(self, index: int):
        cv = C.c_double()
        ok = C.c_int()
        self._chk(self.lib.lbm_scalar_cv(self.h, index, C.byref(cv), C.byref(ok)))
        return cv.value if ok.value else None

    def compartment_cv(self, index: int, impeller_center_z: float | None = None):
        """``qoi::compartment_cv`` (qoi.rs:699-739) of scalar `index`: a list
        of (name, cv or None, cell_count) for top, bulk and near_impeller."""
        cv = np.zeros(3)
        ok = np.zeros(3, np.int32)
        cnt = np.zeros(3, np.int64)
        has = impeller_center_z is not None
        self._chk(self.lib.lbm_scalar_compartment_cv(self.h, index, int(has), float(impeller_center_z or 0.0),
                                                     _dp(cv), _ip(ok), cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return [(name, float(cv[k]) if ok[k] else None, int(cnt[k]))
                for k, name in enumerate(("top", "bulk", "near_impeller"))]


# ---- geometry and QOI helpers ---------------------------------------------------

def radial_impeller(dims, dx, vessel_diameter, liquid_height, diameter, clearance, blades=6, blade_angle_deg=0.0,
                    omega_z=0.0, thickness_fraction=0.04):
    """``build_radial_impeller`` (geometry.rs:815-906): markers in lattice units."""
    lib = _L.lib()
    d = np.array(dims, np.int32)
    cnt = C.c_int64()
    rc = lib.lbm_geom_radial_impeller(_ip(d), dx, vessel_diameter, liquid_height, diameter, clearance, blades,
                                      blade_angle_deg, omega_z, thickness_fraction, None, None, None, None, 0,
                                      C.byref(cnt))
    if rc != 0:
        raise _L.LbmError(rc, "impeller marker generation produced no markers")
    n = cnt.value
    pos = np.zeros((n, 3))
    w = np.zeros(n)
    vel = np.zeros((n, 3))
    c = np.zeros(3)
    rc = lib.lbm_geom_radial_impeller(_ip(d), dx, vessel_diameter, liquid_height, diameter, clearance, blades,
                                      blade_angle_deg, omega_z, thickness_fraction, _dp(pos), _dp(w), _dp(vel),
                                      _dp(c), n, C.byref(cnt))
    if rc != 0:
        raise _L.LbmError(rc, "impeller marker generation failed")
    return pos, w, vel, c


def impeller_surface_markers(dims, dx, vessel_diameter, liquid_height, diameter, clearance, blades=6,
                             blade_angle_deg=0.0, omega_z=0.0, thickness_fraction=0.04, area_per_marker=0.1,
                             seed=20260512):
    """Seeded uniform surface sampling of the same blades (no reference
    counterpart): markers on the blade boxes of build_radial_impeller at one
    marker per `area_per_marker` lattice units^2, each weighted by its area
    (m^2).  Returns (pos, weight, velocity, center) in lattice units."""
    lib = _L.lib()
    d = np.array(dims, np.int32)
    cnt = C.c_int64()
    args = (_ip(d), dx, vessel_diameter, liquid_height, diameter, clearance, blades, blade_angle_deg, omega_z,
            thickness_fraction, area_per_marker, int(seed) & 0xFFFFFFFFFFFFFFFF)
    rc = lib.lbm_geom_impeller_surface(*args, None, None, None, None, 0, C.byref(cnt))
    if rc != 0:
        raise _L.LbmError(rc, "impeller surface sampling produced no markers")
    n = cnt.value
    pos = np.zeros((n, 3))
    w = np.zeros(n)
    vel = np.zeros((n, 3))
    c = np.zeros(3)
    rc = lib.lbm_geom_impeller_surface(*args, _dp(pos), _dp(w), _dp(vel), _dp(c), n, C.byref(cnt))
    if rc != 0:
        raise _L.LbmError(rc, "impeller surface sampling failed")
    return pos, w, vel, c


def power_qois(torque, omega, rho, diameter, volume):
    """``qoi::power_qois`` (qoi.rs:477-505)."""
    out = np.zeros(4)
    rc = _L.lib().lbm_qoi_power(torque, omega, rho, diameter, volume, _dp(out))
    if rc != 0:
        raise ValueError("power QOI inputs must be finite and positive where dimensional")
    return {"power": out[0], "n_hz": out[1], "np": out[2], "p_over_v": out[3]}


def mixing_time_from_cv(t, cv):
    """``qoi::mixing_time_from_cv`` (qoi.rs:534-556)."""
    t = np.ascontiguousarray(t, np.float64)
    cv = np.ascontiguousarray(cv, np.float64)
    out = np.zeros(3)
    rc = _L.lib().lbm_qoi_mixing_time(t.size, _dp(t), _dp(cv), _dp(out))
    if rc != 0:
        return None
    f = lambda v: None if np.isnan(v) else float(v)  # noqa: E731
    return {"cv0": float(out[0]), "t95": f(out[1]), "t99": f(out[2])}


class Batch:
    """Several closed domains stepped together: one persistent collide-stream
    launch over their concatenated tile lists and one set of IBM launches for
    all rotating bodies (lbm_batch_* in include/lbm_hip.h).  The member
    solvers keep their state and stay readable; they must not be changed
    while the batch exists."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        if not self.solvers:
            raise ValueError("a batch needs at least one solver")
        self.ctx = self.solvers[0].ctx
        self.lib = _L.lib()
        arr = (C.c_void_p * len(self.solvers))(*[sv.h for sv in self.solvers])
        h = C.c_void_p()
        self.ctx.check(self.lib.lbm_batch_create(self.ctx.h, arr, len(self.solvers), C.byref(h)))
        self.h = h

    def info(self):
        m, t, c, g = C.c_int(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self.ctx.check(self.lib.lbm_batch_info(self.h, C.byref(m), C.byref(t), C.byref(c), C.byref(g)))
        return {"members": m.value, "tiles": t.value, "forced_cells": c.value, "grid": g.value}

    def set_grid(self, grid: int):
        self.ctx.check(self.lib.lbm_batch_set_grid(self.h, int(grid)))

    def autotune(self, trial_steps: int = 10):
        g, ms = C.c_uint32(), C.c_double()
        self.ctx.check(self.lib.lbm_batch_autotune(self.h, int(trial_steps), C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def run(self, steps: int, graph: bool = True):
        """``steps`` runner steps of every member; returns torque[members, steps, 3]
        (zeros when the members carry no rotating body)."""
        steps = int(steps)
        tq = np.zeros((len(self.solvers), max(steps, 0), 3))
        self.ctx.check(self.lib.lbm_batch_run_span(self.h, steps, int(bool(graph)), _dp(tq)))
        return tq

    def torque_integrals(self, reset: bool = False):
        """The members' running FP64 torque sums over the updates stepped in
        this batch (lbm_batch_torque_integral): (sums[members, 3], updates[members])."""
        out = np.zeros((len(self.solvers), 4))
        self.ctx.check(self.lib.lbm_batch_torque_integral(self.h, int(bool(reset)), _dp(out)))
        return out[:, :3].copy(), out[:, 3].astype(np.int64)

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class WaleLes:
    """``WaleLes`` (les.rs:46-221): WALE SGS-viscosity driver with
    ``Cw = 0.325`` and ``Delta = 1``.  ``update(solver)`` computes ``nu_t``
    from the solver's current velocity gradient and installs
    ``omega_plus = 1/(3 (nu0 + nu_t) + 1/2)`` for the next collisions;
    ``clear(solver)`` removes the field."""

    def __init__(self, tau_eff_max: float | None = None):
        self.cw = WALE_CW
        self.delta = 1.0
        self._tau_eff_max = None
        self.set_tau_eff_max(tau_eff_max)
        self._nu_t = np.zeros(0, np.float32)
        self._diag = {"clipped_cells": 0, "clipped_fraction": 0.0, "max_nu_t_before_clipping": 0.0,
                      "tau_eff_max": None, "nu_t_max": None}

    def with_tau_eff_max(self, tau_eff_max: float) -> "WaleLes":
        self.set_tau_eff_max(tau_eff_max)
        return self

    def set_tau_eff_max(self, tau_eff_max: float | None):
        """les.rs:82-95: the bound must be finite and above 0.5."""
        if tau_eff_max is not None and not (np.isfinite(tau_eff_max) and tau_eff_max > 0.5):
            raise ValueError("tau_eff_max must be finite and greater than 0.5")
        self._tau_eff_max = tau_eff_max

    def tau_eff_max(self):
        return self._tau_eff_max

    def nu_t(self):
        """Last computed eddy viscosity, compact order."""
        return self._nu_t

    def diagnostics(self) -> dict:
        return dict(self._diag)

    def update(self, solver: "Solver"):
        self._diag, self._nu_t = solver.wale_update(self.cw * self.delta, self._tau_eff_max, want_nu_t=True)

    def clear(self, solver: "Solver"):
        solver.set_omega_field(None)


def revolve_schedule(steps: int, snapshots: int):
    """The binomial checkpoint schedule of the torque adjoint as a list of
    (action, argument): 'advance' n, 'save' slot, 'restore' slot, 'adjoint' step."""
    lib = _L.lib()
    n = C.c_int()
    rc = lib.lbm_revolve_schedule(int(steps), int(snapshots), None, 0, C.byref(n))
    if rc != 0:
        raise _L.LbmError(rc, "invalid schedule request")
    buf = np.zeros(2 * n.value, np.int32)
    rc = lib.lbm_revolve_schedule(int(steps), int(snapshots), _ip(buf), n.value, C.byref(n))
    if rc != 0:
        raise _L.LbmError(rc, "schedule")
    names = ["advance", "save", "restore", "adjoint"]
    return [(names[buf[2 * k]], int(buf[2 * k + 1])) for k in range(n.value)]
