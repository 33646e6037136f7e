"""The C-ABI library loads on the CPU and exports every entry point that
include/lbm_hip.h declares; host-only entry points match the oracle."""
import ctypes as C
import os
import re

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_hip.h")
LIB = os.path.join(ROOT, "lbmflow_amd", "liblbm_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 39
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from lbmflow_amd import _lib

    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_host_impeller_generation_matches_oracle():
    import lbmflow_amd as L

    for dims, args in [((256,) * 3, (1.0, 240.0, 240.0, 80.0, 80.0, 6, 0.0, 1.25e-3)),
                       ((96,) * 3, (1 / 96, 1.0, 1.0, 0.5, 0.35, 4, 45.0, 2.0, 0.035))]:
        a = L.radial_impeller(dims, *args)
        kw = {}
        b = O.radial_impeller(dims, *args, inner_kind=0 if args[6] == 0.0 else 1, **kw)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_qoi_entry_points_match_oracle():
    import lbmflow_amd as L

    for args in [(2.0, 2 * np.pi * 3.0, 1000.0, 0.5, 2.0), (-0.3, 1.7, 998.0, 0.1, 0.01)]:
        a = L.power_qois(*args)
        b = O.power_qois(*args)
        for k in ("power", "n_hz", "np", "p_over_v"):
            assert a[k] == b[k]
    t = np.arange(6, dtype=float)
    cv = np.array([np.nan, 0.8, 0.3, 0.04, 0.009, 0.001])
    assert L.mixing_time_from_cv(t, cv) == O.mixing_time(t, cv)
    assert L.mixing_time_from_cv([0.0], [0.0]) is None


def test_context_creation_fails_cleanly_without_a_device():
    import lbmflow_amd._lib as _L

    lib = _L.lib()
    h = C.c_void_p()
    rc = lib.lbm_ctx_create(0, C.byref(h))
    if rc == 0:  # a GPU is present
        lib.lbm_ctx_destroy(h)
    else:
        assert rc in (-1, -3)
