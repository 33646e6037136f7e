"""Scenario ingestion on the GPU: the example bioprocess scenario builds the
same geometry and markers as the oracle's restatement of geometry.rs, and
its runner loop (clear field, IBM, step, scalar CV) matches the oracle
driven with the same mapped parameters: masks, markers and the wall
velocity stamp bit-exact, torque and tracer CV series within 1e-6
relative (the IBM reductions' summation order), Np from the same window.
"""
import json
import math
import os

import numpy as np
import pytest

from lbmflow_amd import scenario as SC
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def example(**edits):
    with open(os.path.join(GOLDEN, "bioprocess_stirred_tank_screening.json")) as fh:
        doc = json.load(fh)
    for k, v in edits.items():
        sec, key = k.split("__")
        doc[sec][key] = v
    return doc


def oracle_geometry(sc, dx, kind):
    r = sc.reactor
    dims = (sc.run["grid_nx"], sc.run["grid_ny"], sc.run["grid_nz"])
    T, H = r["vessel_diameter_m"], r["liquid_height_m"]
    solid = O.tank_mask(dims, dx, T, H)
    b = r["baffles"][0]
    solid, baffle = O.baffle_mask(dims, dx, T, H, b["width_m"], b["thickness_m"], solid)
    imp = r["impellers"][0]
    omega = imp["rotational_speed_rpm"] * 2 * math.pi / 60
    angle, thick, inner = (0.0, 0.04, 0) if kind == "rushton" else (45.0, 0.035, 1)
    pos, w, vel, c = O.radial_impeller(dims, dx, T, H, imp["diameter_m"], imp["clearance_from_bottom_m"],
                                       imp["blade_count"], angle, omega, thick, inner_kind=inner)
    wall = np.zeros((solid.size, 3))
    O.stamp_wall_velocity(dims, pos, vel, wall)
    return solid, baffle, (pos, w, vel, c), wall


@pytest.mark.parametrize("kind,thickness", [("rushton", None), ("pitched_blade", 0.006)])
def test_geometry_matches_oracle(kind, thickness):
    doc = example()
    doc["reactor"]["impellers"][0]["kind"] = kind
    if thickness is not None:
        doc["reactor"]["baffles"][0]["thickness_m"] = thickness
    sc = SC.BioprocessScenario.from_dict(doc)
    geo = sc.build_geometry()
    dx = sc.unit_report_with_diagnostics()["lattice"]["dx_m"]
    solid, baffle, (pos, w, vel, c), wall = oracle_geometry(sc, dx, kind)
    assert np.array_equal(geo.solid, solid)
    assert np.array_equal(geo.baffle_mask, baffle)
    if thickness is None:
        # the example's 2.8 mm baffles are thinner than dx = 2.85 mm and fall
        # between the cell centres at 48^3 (no baffle cell, in the reference too)
        assert geo.baffle_mask.sum() == 0
    else:
        assert geo.baffle_mask.sum() > 100
    imp = geo.impellers[0]
    assert np.array_equal(imp.positions, pos) and np.array_equal(imp.weights, w)
    assert np.array_equal(imp.wall_velocity, vel) and np.array_equal(imp.center, c)
    assert np.array_equal(geo.wall_velocity, wall)


def test_runner_loop_matches_oracle():
    sc = SC.BioprocessScenario.from_dict(example(run__steps=40, outputs__probes_every_n_steps=10))
    run = sc.build_solver()
    rep = run.report
    out = run.run()
    geo = run.geometry
    dims = geo.dims
    dt, dx = rep["lattice"]["dt_s"], rep["lattice"]["dx_m"]
    imp = geo.impellers[0]
    o = O.OracleSolver(19, dims, prec=4, collision=O.TRT, nu=rep["lattice"]["nu_lu"], solid=geo.solid,
                       wall_u=(geo.wall_velocity * dt).astype(np.float32).astype(np.float64))
    m = [p for p in sc.physics if p["kind"] == "passive_scalar"][0]
    k = o.scalar_enable(m["diffusivity_m2_per_s"] * dt / dx ** 2, False)
    pulse = SC.top_hat_pulse(dims, dx, geo.solid, m["initial_pulse"])
    o.scalar_set(k, pulse.astype(np.float32).astype(np.float64))
    assert pulse.sum() > 0
    tq, cv = [], []
    fl = (geo.solid == 0).astype(np.uint8)
    for step in range(1, 41):
        o.clear_force_field()
        d = o.apply_ibm(imp.positions, imp.weights, imp.center, tuple(imp.omega * dt))
        o.step()
        tq.append(-d["torque"][2])
        if step % 10 == 0:
            cv.append(O.scalar_cv(o.scalar_get(k), fl))
    np.testing.assert_allclose(out["torque_lu"], tq, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose([c[1] for c in out["scalar_cv"]], cv, rtol=1e-6)
    assert [c[0] for c in out["scalar_cv"]] == pytest.approx([s * dt for s in (10, 20, 30, 40)])
    rho = sc.fluids["liquid_density_kg_m3"]
    scale = rho * dx ** 5 / dt ** 2
    win = np.array(tq)[np.array([10, 20, 30, 40]) - 1][1:]  # record steps >= 20
    omega = 250 * 2 * math.pi / 60
    q = O.power_qois(win.mean() * scale, omega, rho, 0.0455, sc.reactor["working_volume_m3"])
    assert out["power"]["np"] == pytest.approx(q["np"], rel=1e-6)
    run.close()
