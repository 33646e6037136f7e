"""GPU parity: the HIP library against the oracle on identical inputs.

Bars (reference gate backend_simd_equiv.rs:1-25): TRT/BGK fp32 populations
and moments are bit-exact (the kernels keep the reference's operand order);
central-moment fp32 populations within 1e-6 absolute on the deviation;
flags, bounce-back links and IBM force fields bit-exact; f64 reductions to
summation-order round-off.
"""
import numpy as np
import pytest

import lbmflow_amd as L
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def fluid_pops(f, solid):
    return f[:, solid == 0].astype(np.float64)


def assert_pops_exact(orc, hip, solid, what=""):
    a = fluid_pops(hip.populations(), solid)
    b = fluid_pops(orc.get_f(), solid)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f"{what}: {len(bad)} populations differ, first {bad[:3]}, max |d| {np.abs(a - b).max()}"


def assert_moments_exact(orc, hip, solid, what=""):
    m_h = hip.moments()
    m_o = orc.moments()
    fl = solid == 0
    for k, name in enumerate(("rho", "ux", "uy", "uz")):
        a = m_h[k][fl].astype(np.float64)
        b = m_o[k][fl]
        assert np.array_equal(a, b), f"{what}: moment {name} max |d| {np.abs(a - b).max()}"


def assert_cm_gate(a, b, Q, what=""):
    """Central-moment bar: the reference's own CpuSimd-vs-CpuScalar gate,
    1e-6 absolute on the stored deviations (backend_simd_equiv.rs:1-25), and
    1e-6 relative to the largest physical population f = w + f_dev."""
    _, w, _, _ = O.lattice(Q)
    err = np.abs(a - b).max()
    scale = np.abs(b + np.asarray(w)[:, None]).max()
    assert err <= 1e-6, f"{what}: CM populations max |d| {err}"
    assert err <= 1e-6 * scale, f"{what}: CM populations max |d| {err} > 1e-6 x {scale}"
    return err


def test_context_and_capabilities(gpu_ctx):
    caps = gpu_ctx.capabilities()
    assert caps & 0x1 and caps & 0x2 and caps & 0x40 and caps & 0x80
    assert caps & 0x200 and caps & 0x2000  # open faces, WALE


@pytest.mark.parametrize("Q", [19, 27])
def test_cavity_trt_bitexact(Q):
    n = 18
    solid, wall = cases.cavity_masks(n)
    orc, hip = cases.pair(Q, (n, n, n), collision="trt", nu=0.05 * (n - 2) / 100.0, solid=solid, wall_u=wall)
    assert np.array_equal(hip.solid(), solid)
    cells, qs = hip.boundary_links()
    ec, eq = cases.links_from_solid(solid, (n, n, n), Q)
    assert np.array_equal(cells, ec) and np.array_equal(qs, eq)
    done = 0
    for k in (1, 1, 5, 33):
        orc.step(k)
        hip.run(k)
        done += k
        assert_pops_exact(orc, hip, solid, f"cavity D3Q{Q} step {done}")
        assert_moments_exact(orc, hip, solid, f"cavity D3Q{Q} step {done}")


def test_mask_reset_at_odd_parity_keeps_state():
    """Re-setting the masks mid-run (odd parity: bounced populations sit in
    the fluid cells' own slots) must not disturb the flow."""
    n = 16
    solid, wall = cases.cavity_masks(n)
    orc, hip = cases.pair(27, (n, n, n), collision="trt", nu=0.03, solid=solid, wall_u=wall)
    orc.step(3)
    hip.run(3)
    hip.set_masks(solid, wall)
    assert_pops_exact(orc, hip, solid, "after set_masks at odd parity")
    orc.step(4)
    hip.run(4)
    assert_pops_exact(orc, hip, solid, "4 steps later")


@pytest.mark.parametrize("Q,coll", [(19, "trt"), (27, "trt"), (19, "bgk")])
def test_tgv_periodic_bitexact(Q, coll):
    dims = (16, 12, 10)
    rho, u = cases.tgv_field(dims)
    orc, hip = cases.pair(Q, dims, collision=coll, nu=0.02, periodic=(1, 1, 1))
    orc.init_with(rho, u)
    hip.init_with(rho, u)
    solid = np.zeros(int(np.prod(dims)), np.uint8)
    assert_pops_exact(orc, hip, solid, "init_with")
    for k in (1, 24):
        orc.step(k)
        hip.run(k)
        assert_pops_exact(orc, hip, solid, f"TGV D3Q{Q} {coll}")
    assert_moments_exact(orc, hip, solid, "TGV")


def seam_solids(dims):
    """Solid blocks across the periodic x, y and z seams, an all-solid half
    of the first tile (x 32..63) in some rows, and a partial last tile."""
    nx, ny, nz = dims
    s = np.zeros((nz, ny, nx), np.uint8)
    s[2:5, 3:6, :3] = s[2:5, 3:6, nx - 2:] = 1          # across x = 0 | nx - 1
    s[6:8, :2, 10:20] = s[6:8, ny - 1:, 10:20] = 1      # across y
    s[:2, 5:7, 40:45] = s[nz - 1:, 5:7, 40:45] = 1      # across z
    s[3:6, 1:4, 32:64] = 1                              # an all-solid tile half
    s[4, 4, 66:68] = 1                                  # inside the partial last tile
    return s.reshape(-1)


@pytest.mark.parametrize("coll,exact", [("trt", True), ("cm", False)])
def test_periodic_seam_solids_d3q27(coll, exact):
    """D3Q27 fp32 with solids across the periodic seams: the odd-step fill of
    solid lanes addresses its targets with the wrap, and the bounce-back and
    streaming stay bit-exact (TRT; CM within 1e-6) after odd and even counts."""
    dims = (70, 9, 8)
    solid = seam_solids(dims)
    rho, u = cases.tgv_field(dims, u0=0.03)
    orc, hip = cases.pair(27, dims, collision=coll, nu=0.02, periodic=(1, 1, 1), solid=solid)
    orc.init_with(rho, u)
    hip.init_with(rho, u)
    fl = solid == 0
    for k in (1, 1, 3, 8):
        orc.step(k)
        hip.run(k)
        a = hip.populations()[:, fl].astype(np.float64)
        b = orc.get_f()[:, fl]
        if exact:
            assert np.array_equal(a, b), f"{coll}: max {np.abs(a - b).max()}"
        else:
            assert np.abs(a - b).max() <= 1e-6, f"{coll}: max {np.abs(a - b).max()}"


def test_force_fields_and_stale_moments_bitexact():
    """Non-uniform body force, then clear (moments stay stale, solver.rs:2805-2822)."""
    n = 14
    solid, wall = cases.cavity_masks(n, u_lid=0.03)
    orc, hip = cases.pair(19, (n, n, n), collision="trt", nu=0.04, solid=solid, wall_u=wall,
                          force=(1e-6, 0.0, -2e-6))
    x, y, z = cases.coords((n, n, n))
    ff = np.stack([1e-5 * np.sin(0.5 * x), 2e-5 * np.cos(0.3 * y * z), 0.0 * x], 1).astype(np.float32)
    ff[solid != 0] = 0.0
    orc.set_force_field(ff.astype(np.float64))
    hip.set_body_force_field(ff)
    orc.step(3)
    hip.run(3)
    assert_pops_exact(orc, hip, solid, "with field")
    orc.clear_force_field()
    hip.clear_body_force_field()
    orc.step(4)
    hip.run(4)
    assert_pops_exact(orc, hip, solid, "after clear")
    assert_moments_exact(orc, hip, solid, "after clear")
    uni = np.tile(np.array([[2e-6, -1e-6, 0.0]], np.float32), (n ** 3, 1))
    orc.set_force_field(uni.astype(np.float64))
    hip.set_body_force_field(uni)
    orc.step(2)
    hip.run(2)
    assert_pops_exact(orc, hip, solid, "uniform field")
    assert_moments_exact(orc, hip, solid, "uniform field")


def test_omega_field_bitexact():
    n = 14
    solid, wall = cases.cavity_masks(n)
    orc, hip = cases.pair(27, (n, n, n), collision="trt", nu=0.03, solid=solid, wall_u=wall)
    x, y, z = cases.coords((n, n, n))
    om = (1.2 + 0.3 * np.sin(0.4 * x + 0.2 * z)).astype(np.float32)
    orc.set_omega_field(om.astype(np.float64))
    hip.set_omega_field(om)
    orc.step(12)
    hip.run(12)
    assert_pops_exact(orc, hip, solid, "omega field")


@pytest.mark.parametrize("Q", [19, 27])
@pytest.mark.parametrize("force", [(0.0, 0.0, 0.0), (2e-6, -1e-6, 5e-7)])
def test_central_moment_cavity_within_1e6(Q, force):
    n = 14
    solid, wall = cases.cavity_masks(n, u_lid=0.04)
    x, y, z = cases.coords((n, n, n))
    orc, hip = cases.pair(Q, (n, n, n), collision="cm", nu=0.01, solid=solid, wall_u=wall, force=force,
                          omega_shear=1.7)
    u = np.stack([0.02 * np.sin(0.5 * y), 0.01 * np.cos(0.4 * x), 0.015 * np.sin(0.3 * (x + z))], 1)
    u = u.astype(np.float32).astype(np.float64)
    rho = np.ones(n ** 3)
    orc.init_with(rho, u)
    hip.init_with(rho, u)
    for k in (1, 10, 29):
        orc.step(k)
        hip.run(k)
        assert_cm_gate(fluid_pops(hip.populations(), solid), fluid_pops(orc.get_f(), solid), Q,
                       f"D3Q{Q} cavity, chunk of {k} steps")
    m0 = orc.reduce(1)
    m1 = hip.mass_deviation()
    assert abs(m0 - m1) <= 1e-6 * orc.fluid_cells()
    assert hip.fluid_cells() == orc.fluid_cells()


@pytest.mark.parametrize("Q", [19, 27])
def test_central_moment_periodic_uniform_force(Q):
    dims = (12, 10, 8)
    rho, u = cases.tgv_field(dims, u0=0.03)
    orc, hip = cases.pair(Q, dims, collision="cm", nu=0.005, periodic=(1, 1, 1), force=(1e-5, 0.0, 0.0),
                          omega_shear=1.9)
    orc.init_with(rho, u)
    hip.init_with(rho, u)
    solid = np.zeros(int(np.prod(dims)), np.uint8)
    for k in (7, 33):  # after 7 (odd parity) and 40 (even) steps
        orc.step(k)
        hip.run(k)
        assert_cm_gate(fluid_pops(hip.populations(), solid), fluid_pops(orc.get_f(), solid), Q,
                       f"D3Q{Q} periodic, chunk of {k} steps")
    mo = np.array([orc.reduce(2, a) for a in range(3)])
    mh = hip.momentum()
    assert np.abs(mo - mh).max() <= 1e-6 * np.prod(dims)


def test_reductions_match_oracle():
    n = 16
    solid, wall = cases.cavity_masks(n)
    orc, hip = cases.pair(19, (n, n, n), collision="trt", nu=0.02, solid=solid, wall_u=wall)
    orc.step(20)
    hip.run(20)
    assert hip.fluid_cells() == orc.fluid_cells() == (n - 2) ** 3
    assert abs(hip.mass_deviation() - orc.reduce(1)) <= 1e-12 * n ** 3
    for a in range(3):
        assert abs(hip.reduce(2, a) - orc.reduce(2, a)) <= 1e-12 * n ** 3


@pytest.mark.parametrize("storage,tol", [("f16", 3e-3), ("bf16", 2e-2)])
def test_16bit_storage_tracks_f32(storage, tol):
    n = 16
    solid, wall = cases.cavity_masks(n, u_lid=0.05)
    import lbmflow_amd as L

    ref = L.Solver(19, (n, n, n), nu=0.02, solid=solid, wall_u=wall)
    low = L.Solver(19, (n, n, n), nu=0.02, solid=solid, wall_u=wall, storage=storage)
    ref.run(200)
    low.run(200)
    _, ux_r, uy_r, uz_r = ref.moments()
    _, ux_l, uy_l, uz_l = low.moments()
    fl = solid == 0
    ur = np.sqrt(ux_r ** 2 + uy_r ** 2 + uz_r ** 2)[fl]
    ul = np.sqrt(ux_l ** 2 + uy_l ** 2 + uz_l ** 2)[fl]
    rel = np.linalg.norm(ur - ul) / np.linalg.norm(ur)
    assert rel < tol, f"{storage} velocity L2rel {rel}"


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_stage_roundtrip(storage):
    n = 10
    solid, wall = cases.cavity_masks(n)
    hip = L.Solver(27, (n, n, n), nu=0.02, solid=solid, wall_u=wall, storage=storage)
    rng = np.random.default_rng(7)
    _, w, _, _ = O.lattice(27)
    f = (w[:, None] * 0.05 * rng.standard_normal((27, n ** 3))).astype(np.float32)
    f[:, solid != 0] = 0.0
    hip.set_populations(f)
    g = hip.populations()
    if storage == "f32":
        assert np.array_equal(f, g)
    else:
        rel = np.abs(f - g).max() / np.abs(f).max()
        assert rel < (1e-3 if storage == "f16" else 8e-3)
    hip.run(3)  # odd parity
    g3 = hip.populations()
    hip.set_populations(g3)
    assert np.array_equal(hip.populations(), g3) or storage != "f32"


def test_tank_flags_and_links_bitexact():
    n = 40
    solid, (T, H, bw, bt) = cases.tank_geometry(n, T=36.0, H=34.0)
    hip2 = L.Solver(27, (n, n, n), nu=0.01, collision="cm")
    with pytest.raises(L.LbmError):
        hip2.run(1)  # all-fluid box with closed (non-periodic) faces: open faces are not on this path
    hip2.build_stirred_tank(1.0, T, H, 4, bw, bt)
    hip2.run(1)
    assert np.array_equal(hip2.solid(), solid)
    cells, qs = hip2.boundary_links()
    ec, eq = cases.links_from_solid(solid, (n, n, n), 27)
    assert np.array_equal(cells, ec) and np.array_equal(qs, eq)


def _tank_ibm_case(n=32, Q=19, collision="trt", steps=6):
    solid, (T, H, bw, bt) = cases.tank_geometry(n, T=30.0, H=30.0)
    D = T / 3.0
    omega_z = 0.05 / (0.5 * D)
    pos, w, vel, c = O.radial_impeller((n, n, n), 1.0, T, H, D, T / 3.0, 6, 0.0, omega_z)
    orc, hip = cases.pair(Q, (n, n, n), collision=collision, nu=0.02, solid=solid)
    hip.set_rotating_body(pos, w, c, (0.0, 0.0, omega_z))
    return solid, pos, w, c, omega_z, orc, hip


def test_rotating_ibm_force_field_bitexact():
    solid, pos, w, c, omega_z, orc, hip = _tank_ibm_case()
    for step in range(6):
        orc.clear_force_field()
        hip.clear_body_force_field()
        do = orc.apply_ibm(pos, w, c, (0.0, 0.0, omega_z))
        dh = hip.apply_rotating_ibm()
        fo = orc.force_field()
        fh = hip.body_force_field()
        assert np.array_equal(fh.astype(np.float64), fo), f"force field differs at step {step}"
        assert dh["iterations"] == do["iterations"]
        np.testing.assert_allclose(dh["torque"], do["torque"], rtol=1e-10, atol=1e-15)
        np.testing.assert_allclose(dh["marker_force"], do["marker_force"], rtol=1e-10, atol=1e-15)
        assert abs(dh["slip_max"] - do["slip_max"]) <= 1e-12 * max(1.0, do["slip_max"])
        orc.step()
        hip.step()
        assert_pops_exact(orc, hip, solid, f"IBM tank step {step}")


@pytest.mark.parametrize("Q", [19, 27])
def test_rotating_ibm_cm(Q):
    solid, pos, w, c, omega_z, orc, hip = _tank_ibm_case(n=28, Q=Q, collision="cm")
    for step in range(5):
        orc.clear_force_field()
        hip.clear_body_force_field()
        do = orc.apply_ibm(pos, w, c, (0.0, 0.0, omega_z))
        dh = hip.apply_rotating_ibm()
        # the in-plane components are round-off around zero (6-fold symmetric body)
        np.testing.assert_allclose(dh["torque"], do["torque"], rtol=1e-5, atol=1e-7 * abs(do["torque"][2]))
        orc.step()
        hip.step()
    assert_cm_gate(fluid_pops(hip.populations(), solid), fluid_pops(orc.get_f(), solid), Q, f"D3Q{Q} IBM tank")


@pytest.mark.parametrize("lattice", ["hydro", "d3q7"])
def test_scalar_ade_bitexact(lattice):
    n = 16
    solid, wall = cases.cavity_masks(n, u_lid=0.05)
    orc, hip = cases.pair(19, (n, n, n), collision="trt", nu=0.02, solid=solid, wall_u=wall)
    x, y, z = cases.coords((n, n, n))
    c0 = np.exp(-((x - 5.0) ** 2 + (y - 9.0) ** 2 + (z - 10.0) ** 2) / (2 * 2.0 ** 2))
    c0[solid != 0] = 0.0
    c0 = c0.astype(np.float32).astype(np.float64)
    D = 0.02
    orc.step(5)
    hip.run(5)
    ko = orc.scalar_enable(D, lattice == "d3q7")
    kh = hip.enable_scalar_distribution(D, lattice)
    orc.scalar_set(ko, c0)
    hip.set_scalar_concentration(kh, c0)
    for k in (1, 19):
        orc.step(k)
        hip.run(k)
        a = hip.gather_scalar_concentration(kh).astype(np.float64)
        b = orc.scalar_get(ko)
        assert np.array_equal(a, b), f"scalar {lattice}: max |d| {np.abs(a - b).max()}"
    cv_o = O.scalar_cv(b, (solid == 0).astype(np.uint8))
    cv_h = hip.scalar_cv(kh)
    assert abs(cv_h - cv_o) <= 1e-12 * cv_o
