"""ctypes binding of the C ABI declared in ``include/lbm_hip.h``.

The library is the product path: if it is missing, importing the solver
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LBMFLOW_LIB") or os.path.join(HERE, "liblbm_hip.so")

LBM_OK = 0
ERRORS = {-1: "invalid argument", -2: "unsupported", -3: "device error", -4: "out of device memory",
          -5: "call out of order"}

STORAGE = {"f32": 0, "f16": 1, "bf16": 2}
COLLISION = {"bgk": 0, "trt": 1, "cm": 2}


class LbmError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class StepParams(C.Structure):
    _fields_ = [("collision", C.c_int), ("omega_p", C.c_double), ("omega_m", C.c_double),
                ("omega_shear", C.c_double), ("force", C.c_float * 3)]


class IbmConfig(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("slip_tolerance", C.c_double), ("kernel_radius", C.c_int),
                ("relaxation", C.c_double)]


class IbmDiag(C.Structure):
    _fields_ = [("iterations", C.c_int), ("slip_max", C.c_double), ("slip_rms", C.c_double),
                ("slip_max_rel", C.c_double), ("slip_rms_rel", C.c_double), ("torque", C.c_double * 3),
                ("fluid_force", C.c_double * 3), ("marker_force", C.c_double * 3),
                ("momentum_error_rel", C.c_double)]


class CkptInfo(C.Structure):
    _fields_ = [("bytes", C.c_uint64), ("payload_hash", C.c_uint64), ("mask_hash", C.c_uint64),
                ("step", C.c_int64)]


class VolumeSource(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3), ("kind", C.c_int), ("q_lu", C.c_double),
                ("u", C.c_double * 3)]


class FacePatch(C.Structure):
    _fields_ = [("face", C.c_int), ("lo", C.c_int * 2), ("hi", C.c_int * 2), ("kind", C.c_int),
                ("value", C.c_double * 3)]


class WaleDiag(C.Structure):
    _fields_ = [("clipped_cells", C.c_int64), ("clipped_fraction", C.c_double),
                ("max_nu_t_before_clipping", C.c_double), ("tau_eff_max", C.c_double), ("nu_t_max", C.c_double)]


_lib = None

# name -> (restype, argtypes)
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_pI = C.POINTER(C.c_int)
_pD = C.POINTER(C.c_double)
_pF = C.POINTER(C.c_float)
_pU8 = C.POINTER(C.c_uint8)
_pI64 = C.POINTER(C.c_int64)
_pU32 = C.POINTER(C.c_uint32)

SIGNATURES = {
    "lbm_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "lbm_ctx_destroy": (None, [_P]),
    "lbm_last_error": (C.c_char_p, [_P]),
    "lbm_capabilities": (_I, [_P, _pU32]),
    "lbm_build_info": (C.c_char_p, []),
    "lbm_stream_copy": (_I, [_P, C.c_size_t, _I, _pD]),
    "lbm_access_copy": (_I, [_P, _I, _I, C.c_size_t, _I, _pD]),
    "lbm_fields_alloc": (_I, [_P, _I, _I, _pI, _pI, C.POINTER(_P)]),
    "lbm_fields_free": (None, [_P]),
    "lbm_fields_device_bytes": (C.c_size_t, [_P]),
    "lbm_set_masks": (_I, [_P, _pU8, _pF]),
    "lbm_read_solid": (_I, [_P, _pU8]),
    "lbm_boundary_links": (_I, [_P, _pI64, _pI64, _pU8, C.c_int64]),
    "lbm_set_params": (_I, [_P, C.POINTER(StepParams)]),
    "lbm_set_omega_field": (_I, [_P, _pF]),
    "lbm_wale_update": (_I, [_P, _D, _D, _pF, C.POINTER(WaleDiag)]),
    "lbm_set_wale": (_I, [_P, _I, _D, _D]),
    "lbm_set_face_bc": (_I, [_P, _I, _I, _pD]),
    "lbm_set_force_probe": (_I, [_P, C.POINTER(C.c_uint8)]),
    "lbm_set_volume_sources": (_I, [_P, _I, C.POINTER(VolumeSource)]),
    "lbm_set_face_patches": (_I, [_P, _I, C.POINTER(FacePatch)]),
    "lbm_set_inlet_profile": (_I, [_P, _I, _pD]),
    "lbm_probed_force": (_I, [_P, _pD]),
    "lbm_gather_omega_field": (_I, [_P, _pF, _pI]),
    "lbm_stage_in": (_I, [_P, _pF]),
    "lbm_stage_out": (_I, [_P, _pF]),
    "lbm_init_with": (_I, [_P, _pF, _pF]),
    "lbm_update_moments": (_I, [_P]),
    "lbm_set_force_field": (_I, [_P, _pF]),
    "lbm_clear_force_field": (_I, [_P]),
    "lbm_read_force_field": (_I, [_P, _pF, _pI]),
    "lbm_run_span": (_I, [_P, _I]),
    "lbm_synchronize": (_I, [_P]),
    "lbm_kernel_timing": (_I, [_P, _I]),
    "lbm_kernel_timing_read": (_I, [_P, _pD, _pI64]),
    "lbm_force_list_size": (C.c_int64, [_P]),
    "lbm_time": (C.c_int64, [_P]),
    "lbm_read_moments": (_I, [_P, _pF, _pF, _pF, _pF]),
    "lbm_read_moments_noalloc": (_I, [_P, _pF, _pF, _pF, _pF]),
    "lbm_reduce": (_I, [_P, _I, _I, _pD]),
    "lbm_geom_wall_rims": (_I, [_P, _pI, _pD]),
    "lbm_geom_stirred_tank": (_I, [_P, _D, _D, _D, _I, _D, _D]),
    "lbm_geom_radial_impeller": (_I, [_pI, _D, _D, _D, _D, _D, _I, _D, _D, _D, _pD, _pD, _pD, _pD, C.c_int64,
                                      _pI64]),
    "lbm_geom_impeller_surface": (_I, [_pI, _D, _D, _D, _D, _D, _I, _D, _D, _D, _D, C.c_uint64, _pD, _pD, _pD, _pD,
                                       C.c_int64, _pI64]),
    "lbm_ibm_set_body": (_I, [_P, C.c_int64, _pD, _pD, _pD, _pD, C.POINTER(IbmConfig)]),
    "lbm_ibm_apply": (_I, [_P, C.POINTER(IbmDiag)]),
    "lbm_run_ibm_span": (_I, [_P, _I, _I, _pD, C.POINTER(IbmDiag)]),
    "lbm_ibm_torque_integral": (_I, [_P, _I, _pD]),
    "lbm_autotune_tiles": (_I, [_P, _I, _pU32, _pD]),
    "lbm_batch_create": (_I, [_P, C.POINTER(_P), _I, C.POINTER(_P)]),
    "lbm_batch_destroy": (None, [_P]),
    "lbm_batch_info": (_I, [_P, _pI, _pU32, _pU32, _pU32]),
    "lbm_batch_set_grid": (_I, [_P, C.c_uint32]),
    "lbm_batch_run_span": (_I, [_P, _I, _I, _pD]),
    "lbm_batch_autotune": (_I, [_P, _I, _pU32, _pD]),
    "lbm_batch_torque_integral": (_I, [_P, _I, _pD]),
    "lbm_adjoint_torque": (_I, [_P, _I, _I, _I, _pD, _pD, _pD, C.POINTER(C.c_int64)]),
    "lbm_revolve_schedule": (_I, [_I, _I, _pI, _I, _pI]),
    "lbm_checkpoint_write_part": (_I, [_P, C.c_char_p, C.POINTER(CkptInfo)]),
    "lbm_checkpoint_read_part": (_I, [_P, C.c_char_p, C.c_int64, C.POINTER(CkptInfo)]),
    "lbm_fnv1a64": (C.c_uint64, [C.c_void_p, C.c_size_t]),
    "lbm_scalar_enable": (_I, [_P, _D, _I, _pI]),
    "lbm_scalar_set": (_I, [_P, _I, _pF]),
    "lbm_scalar_gather": (_I, [_P, _I, _pF]),
    "lbm_scalar_cv": (_I, [_P, _I, _pD, _pI]),
    "lbm_scalar_compartment_cv": (_I, [_P, _I, _I, _D, _pD, _pI, _pI64]),
    "lbm_qoi_power": (_I, [_D, _D, _D, _D, _D, _pD]),
    "lbm_qoi_mixing_time": (_I, [C.c_int64, _pD, _pD, _pD]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run lbmflow_amd.build.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
