"""lbmflow_amd — MI355X (gfx950) lattice-Boltzmann hot path of lbm-core.

The compute runs in ``liblbm_hip.so`` (hand-written HIP kernels behind the C
ABI in ``include/lbm_hip.h``); this package is the host-side mirror of the
reference's Solver/Backend interface over that ABI.
"""
from .solver import (  # noqa: F401
    MAGIC_STD,
    WALE_CW,
    Batch,
    Context,
    DirectForcingConfig,
    Solver,
    default_context,
    impeller_surface_markers,
    mixing_time_from_cv,
    omegas,
    power_qois,
    radial_impeller,
    revolve_schedule,
    WaleLes,
)
from ._lib import LbmError, lib  # noqa: F401

__all__ = ["Solver", "Batch", "Context", "DirectForcingConfig", "omegas", "radial_impeller", "power_qois",
           "mixing_time_from_cv", "MAGIC_STD", "LbmError", "lib", "default_context", "revolve_schedule",
           "WaleLes", "WALE_CW", "impeller_surface_markers"]
