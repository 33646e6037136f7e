"""lbm_read_moments_noalloc (staging in the library's fixed device array)
against lbm_read_moments (per-call staging allocation), through the C ABI:
bit-identical on two calls in a row, the second on a smaller domain, and a
domain above the staging maximum refused with LBM_ERR_INVALID."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_CELLS = 4194304  # LBM_NOALLOC_MAX_CELLS in include/lbm_hip.h


def _read(s, name):
    n = s.n
    out = [np.full(n, np.nan, np.float32) for _ in range(4)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_float)) for a in out]
    rc = getattr(s.lib, name)(s.h, *ptrs)
    return rc, np.stack(out)


def _cavity(Q, dims, storage, steps):
    import lbmflow_amd as L
    from oracle import oracle as O

    u = np.zeros((6, 3))
    u[5] = [0.05, 0.0, 0.0]
    solid, wall = O.build_wall_rims(dims, [1] * 6, u)
    s = L.Solver(Q, dims, nu=0.02, collision="trt", storage=storage, solid=solid,
                 wall_u=wall.astype(np.float32).astype(np.float64))
    rng = np.random.default_rng(7)
    ff = np.zeros((s.n, 3), np.float32)
    fl = np.flatnonzero(solid.reshape(-1) == 0)
    pick = rng.choice(fl, size=min(40, fl.size), replace=False)
    ff[pick] = rng.normal(0.0, 1e-4, (pick.size, 3)).astype(np.float32)
    s.set_body_force_field(ff)
    s.run(steps)
    return s


def test_noalloc_moments_match_the_allocating_entry_point():
    cases = [(19, (24, 20, 18), "f32", 7), (27, (12, 10, 8), "f16", 4)]
    for Q, dims, storage, steps in cases:
        print(f"[no_alloc] D3Q{Q} {dims} {storage} after {steps} steps", flush=True)
        s = _cavity(Q, dims, storage, steps)
        rc_a, a = _read(s, "lbm_read_moments")
        assert rc_a == 0
        rc_b, b = _read(s, "lbm_read_moments_noalloc")
        assert rc_b == 0, s.lib.lbm_last_error(s.ctx.h)
        assert np.isfinite(a).all()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        # a second call on the same state reads the same again
        rc_c, c = _read(s, "lbm_read_moments_noalloc")
        assert rc_c == 0 and np.array_equal(b.view(np.uint32), c.view(np.uint32))
        s.close()


def test_noalloc_moments_refuse_domains_above_the_staging_maximum():
    import lbmflow_amd as L

    dims = (2049, 2048, 1)
    assert dims[0] * dims[1] * dims[2] > MAX_CELLS
    print(f"[no_alloc] refusal {dims}", flush=True)
    s = L.Solver(19, dims, nu=0.02, collision="trt", periodic=(True, True, True))
    rc, _ = _read(s, "lbm_read_moments_noalloc")
    assert rc == -1
    assert b"staging maximum" in s.lib.lbm_last_error(s.ctx.h)
    s.close()
