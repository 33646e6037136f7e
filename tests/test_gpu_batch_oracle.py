"""The batched sweep of stirred tanks (64 tanks of the sweep workload, seed
20260512: edges U[48,160], tip speed U[0.02,0.06], nu U[0.002,0.02], D3Q19
TRT fp32) against the oracle, member by member: four of the sweep's tanks
(the smallest, edge 51; the largest, edge 159; and the first two drawn) run
together in one persistent launch with the autotuned grid and the
graph-captured step, and each must match the oracle running that tank alone
as the reference's sweep does (crates/lbm-cli/src/sweep.rs:61-67): flags and
bounce-back links bit-exact, populations bit-exact (TRT keeps the
reference's operand order), per-step torque and the device's running torque
integral within 1e-10 relative, and the power number of the last-half window (runner.rs:1514-1546) within 0.5%."""
import numpy as np
import pytest

import bench
import cases
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED = 20260512
PICK = (24, 28, 0, 1)  # indices into the sweep's 64 draws


def sweep_specs(seed=SEED, count=64):
    rng = bench.splitmix64(seed)
    return [(48 + int(next(rng) * 113), 0.02 + 0.04 * next(rng), 0.002 + 0.018 * next(rng)) for _ in range(count)]


def oracle_member(n, tip, nu):
    T = 15.0 * n / 16.0
    D = T / 3.0
    omega = tip / (0.5 * D)
    solid, _ = cases.tank_geometry(n, T=T, H=T, baffle_width=T / 10.0, baffle_thickness=T / 48.0)
    pos, w, vel, c = O.radial_impeller((n, n, n), 1.0, T, T, D, T / 3.0, 6, 0.0, omega)
    orc = O.OracleSolver(19, (n, n, n), prec=4, collision=O.TRT, nu=nu, solid=solid)
    return orc, solid, pos, w, c, omega, D


def test_sweep_members_match_the_oracle():
    import lbmflow_amd as L

    specs = sweep_specs()
    picked = [specs[k] for k in PICK]
    ns = [sp[0] for sp in picked]
    assert min(ns) == min(sp[0] for sp in specs) and max(ns) >= 140
    print(f"[batch-oracle] members {picked}", flush=True)
    members, meta = [], []
    for n, tip, nu in picked:
        s, m = bench.small_tank(n, tip, nu)
        members.append(s)
        meta.append(m)
    batch = L.Batch(members)
    grid, _ = batch.autotune(4)
    assert grid > 0
    steps = 20
    tq = batch.run(steps, graph=True)
    assert tq.shape == (len(picked), steps, 3)
    tsum, tcount = batch.torque_integrals()
    assert (tcount == steps).all()
    for k, ((n, tip, nu), s) in enumerate(zip(picked, members)):
        print(f"[batch-oracle] oracle tank {n}^3", flush=True)
        orc, solid, pos, w, c, omega, D = oracle_member(n, tip, nu)
        assert np.array_equal(s.solid(), solid), f"member {k}: solid flags"
        cells, qs = s.boundary_links()
        ec, eq = cases.links_from_solid(solid, (n, n, n), 19)
        assert np.array_equal(cells, ec) and np.array_equal(qs, eq), f"member {k}: links"
        t_o = []
        for _ in range(steps):
            orc.clear_force_field()
            d = orc.apply_ibm(pos, w, c, (0.0, 0.0, omega))
            orc.step()
            t_o.append(d["torque"])
        t_o = np.array(t_o)
        np.testing.assert_allclose(tq[k], t_o, rtol=1e-10, atol=1e-12 * np.abs(t_o).max(),
                                   err_msg=f"member {k}: torque")
        # the device's running FP64 torque integral against the oracle's sum
        np.testing.assert_allclose(tsum[k], t_o.sum(axis=0), rtol=1e-10, atol=1e-12 * np.abs(t_o).sum(),
                                   err_msg=f"member {k}: torque integral")
        fl = solid == 0
        a = s.populations()[:, fl].astype(np.float64)
        b = orc.get_f()[:, fl]
        assert np.array_equal(a, b), f"member {k}: populations max |d| {np.abs(a - b).max()}"
        win = np.arange(1, steps + 1) >= steps // 2
        np_g = L.power_qois(float(np.mean(-tq[k, win, 2])), omega, 1.0, D, 1.0)["np"]
        np_o = L.power_qois(float(np.mean(-t_o[win, 2])), omega, 1.0, D, 1.0)["np"]
        assert abs(np_g - np_o) <= 5e-3 * abs(np_o), f"member {k}: Np {np_g} vs {np_o}"
    batch.close()
