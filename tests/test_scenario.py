"""Scenario ingestion on the CPU: parsing, refusals and the unit report.

The cases and their expected values are the reference's own tests
(crates/lbm-scenario/tests/bioprocess_units.rs and the unit tests of
crates/lbm-scenario/src/bioprocess.rs:1425-1713) on the same JSON, plus
the shipped example scenario (examples/bioprocess/stirred_tank_screening.json,
kept as tests/golden/bioprocess_stirred_tank_screening.json).  No GPU call
is made: every check here fails before a device is needed.
"""
import copy
import json
import math
import os

import pytest

from lbmflow_amd import scenario as SC

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def base_scenario():
    """bioprocess_units.rs:8-71."""
    return {
        "version": "bioprocess-1.0",
        "name": "bcfd-004-units",
        "credibility_tier": "screening",
        "reactor": {
            "kind": "stirred_tank", "vessel_diameter_m": 0.2, "liquid_height_m": 0.2, "working_volume_m3": 0.006,
            "impellers": [{"kind": "rushton", "diameter_m": 0.05, "clearance_from_bottom_m": 0.06,
                           "rotational_speed_rpm": 60.0, "blade_count": 6}],
            "baffles": [], "spargers": [],
        },
        "fluids": {"liquid_density_kg_m3": 1000.0, "liquid_viscosity_pa_s": 0.1, "gas_density_kg_m3": None,
                   "gas_viscosity_pa_s": None, "surface_tension_n_m": None, "oxygen_diffusivity_m2_per_s": None,
                   "henry_constant": None},
        "operation": {"duration_s": 30.0, "gas_inlet_temp_c": None, "initial_condition": {"kind": "quiescent"}},
        "physics": [{"kind": "single_phase"}],
        "qoi": {"power": {}, "mixing": None, "gas_holdup": None, "bubble_size": None, "kla": None,
                "shear_exposure": None, "oxygen_exposure": None, "calibration_dataset_id": None,
                "holdout_dataset_id": None},
        "run": {"steps": 100, "dt_s": 0.0008, "grid_nx": 100, "grid_ny": 100, "grid_nz": 100, "backend": "cpu",
                "precision": "f64", "lattice": "d3q19"},
        "outputs": {"manifest_path": "out/bcfd-004-units/manifest.json", "fields_every_n_steps": None,
                    "probes_every_n_steps": None, "emit_qoi_json": True, "emit_qoi_csv": False},
    }


def parse(doc):
    return SC.BioprocessScenario.from_json_str(json.dumps(doc))


def codes(report, rejections):
    return [i["code"] for i in report["feasibility"]["rejections" if rejections else "warnings"]]


def assert_close(actual, expected):  # bioprocess_units.rs:84-90
    scale = max(abs(actual), abs(expected), 1.0)
    assert abs(actual - expected) <= 1.0e-12 * scale


PHASE = {"kind": "resolved_phase_field", "interface_width_m": 0.01, "mobility_m2_per_s": 1.0e-5,
         "contact_angle_deg": None}


def test_reynolds_matches_analytical_impeller():
    assert_close(parse(base_scenario()).compute_unit_report()["groups"]["reynolds"], 25.0)


def test_optional_groups_absent_without_surface_tension():
    d = base_scenario()
    d["fluids"]["gas_density_kg_m3"] = 1.2
    d["physics"] = [PHASE]
    rep = parse(d).unit_report_with_diagnostics()
    assert rep["groups"]["weber"] is None and rep["groups"]["eotvos"] is None and rep["groups"]["morton"] is None
    assert "WEBER_UNAVAILABLE" in codes(rep, False)


def test_mach_lattice_warn_at_0_11_reject_at_0_31():
    warn = base_scenario()
    warn["run"]["dt_s"] = 0.0025403411844343537
    rep = parse(warn).compute_unit_report()
    assert rep["groups"]["mach_lattice"] > SC.MACH_LATTICE_WARN_THRESHOLD
    assert "MA_LATTICE_HIGH" in codes(rep, False)
    rej = base_scenario()
    rej["run"]["dt_s"] = 0.007159143792496811
    sc = parse(rej)
    raw = sc.unit_report_with_diagnostics()
    assert raw["groups"]["mach_lattice"] > SC.MACH_LATTICE_REJECT_THRESHOLD
    assert "MA_LATTICE_TOO_HIGH" in codes(raw, True)
    with pytest.raises(SC.ScenarioError) as e:
        sc.compute_unit_report()
    assert e.value.kind == "out_of_validity_range"
    # the chosen dt put the Mach number just past each threshold
    assert raw["groups"]["mach_lattice"] == pytest.approx(0.31, rel=1e-6)
    assert rep["groups"]["mach_lattice"] == pytest.approx(0.11, rel=1e-6)


def test_tau_near_half_warn_and_below_half_rejects():
    warn = base_scenario()
    warn["run"]["dt_s"] = 0.0002
    rep = parse(warn).compute_unit_report()
    assert rep["lattice"]["tau_lu"] <= SC.TAU_NEAR_HALF_WARN_THRESHOLD
    assert "TAU_NEAR_HALF" in codes(rep, False)
    bad = base_scenario()
    bad["fluids"]["liquid_viscosity_pa_s"] = 0.0
    with pytest.raises(SC.ScenarioError) as e:
        parse(bad).compute_unit_report()
    assert e.value.kind == "out_of_validity_range"


def test_cahn_thin_and_thick_reject_only_with_phase_field():
    assert parse(base_scenario()).compute_unit_report()["groups"]["cahn"] is None
    thin = base_scenario()
    thin["physics"] = [dict(PHASE, interface_width_m=0.001)]
    assert "INTERFACE_WIDTH_TOO_THIN" in codes(parse(thin).unit_report_with_diagnostics(), True)
    thick = base_scenario()
    thick["physics"] = [dict(PHASE, interface_width_m=0.12)]
    assert "INTERFACE_WIDTH_TOO_LARGE" in codes(parse(thick).unit_report_with_diagnostics(), True)


def test_bubble_under_resolved_rejected():
    d = base_scenario()
    d["reactor"]["spargers"] = [{"kind": "ring", "center_z_m": 0.02, "outer_radius_m": 0.03, "orifice_count": 8,
                                 "orifice_diameter_m": 0.001, "gas_volumetric_flow_m3_per_s": 1.0e-6, "vvm": None,
                                 "inlet_phase": "gas"}]
    d["fluids"]["gas_density_kg_m3"] = 1.2
    d["fluids"]["surface_tension_n_m"] = 0.072
    d["physics"] = [PHASE]
    assert "BUBBLE_UNDER_RESOLVED" in codes(parse(d).unit_report_with_diagnostics(), True)


def test_grid_reynolds_warning():
    d = base_scenario()
    d["fluids"]["liquid_viscosity_pa_s"] = 0.001
    rep = parse(d).unit_report_with_diagnostics()
    assert rep["lattice"]["u_ref_lu"] / rep["lattice"]["nu_lu"] > SC.BIOPROCESS_GRID_REYNOLDS_WARN_THRESHOLD
    assert "GRID_REYNOLDS_HIGH" in codes(rep, False)


def test_matching_priority_declares_only_re_as_mandatory():
    mp = parse(base_scenario()).compute_unit_report()["matching_priority"]
    assert mp["achieved"] == ["Re"]
    assert any("Sc/Pe/Da" in m for m in mp["missing"])


def test_unit_report_serialises_without_nan():
    rep = parse(base_scenario()).compute_unit_report()
    text = json.dumps(rep)
    assert "NaN" not in text and json.loads(text) == rep


def test_unit_mapping_definitions():
    """units.rs:167-176 on the base case: dx = T / nx, u_ref = N D,
    nu = mu / rho, lattice conversions by dt/dx and dt/dx^2."""
    rep = parse(base_scenario()).compute_unit_report()["lattice"]
    dx = 0.2 / 100
    assert_close(rep["dx_m"], dx)
    assert_close(rep["u_ref_lu"], 1.0 * 0.05 * 0.0008 / dx)
    assert_close(rep["nu_lu"], 1e-4 * 0.0008 / dx ** 2)
    assert_close(rep["tau_lu"], 3 * rep["nu_lu"] + 0.5)


# ---- refusals (bioprocess.rs:34-133, 288-337, 1515-1713) -------------------------

def _err(doc):
    with pytest.raises(SC.ScenarioError) as e:
        parse(doc)
    return e.value


def test_unknown_field_is_refused():
    d = base_scenario()
    d["reactor"]["colour"] = "blue"
    assert "unknown field" in _err(d).message
    d = base_scenario()
    d["surprise"] = 1
    assert "unknown field" in _err(d).message


def test_dished_bottom_not_implemented():
    d = base_scenario()
    d["reactor"]["bottom"] = "dished"
    e = _err(d)
    assert e.kind == "not_implemented" and "dished" in e.message


def test_d2q9_and_small_grids_rejected():
    d = base_scenario()
    d["run"]["lattice"] = "d2q9"
    assert _err(d).kind == "out_of_validity_range"
    d = base_scenario()
    d["run"]["grid_nz"] = 2
    assert "1-cell solid rim" in _err(d).message


def test_spargers_need_a_gas_model_and_gas_inlet():
    sp = {"kind": "ring", "center_z_m": 0.02, "outer_radius_m": 0.03, "orifice_count": 8,
          "orifice_diameter_m": 0.004, "gas_volumetric_flow_m3_per_s": 1e-6, "inlet_phase": "gas"}
    d = base_scenario()
    d["reactor"]["spargers"] = [sp]
    e = _err(d)
    assert e.kind == "missing_dependency" and e.reason["depends_on"] == "gas_model_in_physics"
    d["physics"] = [PHASE]
    d["reactor"]["spargers"] = [dict(sp, inlet_phase="liquid")]
    assert "inlet phase must be gas" in _err(d).message
    d["reactor"]["spargers"] = [dict(sp, raw_phi_boundary_fields=["phi"])]
    assert "raw phi boundary fields" in _err(d).message


def test_evidence_tier_needs_datasets_and_distinct_ids():
    d = base_scenario()
    d["credibility_tier"] = "evidence"
    e = _err(d)
    assert e.kind == "evidence_gate_failed"
    d = base_scenario()
    d["qoi"]["calibration_dataset_id"] = d["qoi"]["holdout_dataset_id"] = "same"
    assert "calibration and holdout" in _err(d).message


def test_kla_requires_oxygen():
    d = base_scenario()
    d["qoi"]["kla"] = {}
    e = _err(d)
    assert e.kind == "missing_dependency" and e.reason["depends_on"] == "oxygen"


def test_geometry_resolution_floor_by_tier():
    """bioprocess.rs:1515-1526: D = T/3 on a 48^3 grid (16 cells across the
    impeller) passes the screening floor of 12 cells, not the engineering
    floor of 32."""
    d = base_scenario()
    d["reactor"].update(vessel_diameter_m=1.0, liquid_height_m=1.0, working_volume_m3=0.78539816339)
    d["reactor"]["impellers"][0].update(diameter_m=1.0 / 3.0, clearance_from_bottom_m=0.35,
                                        rotational_speed_rpm=120.0)
    d["run"].update(grid_nx=48, grid_ny=48, grid_nz=48)
    sc = parse(d)
    dx = sc.unit_report_with_diagnostics()["lattice"]["dx_m"]
    sc.validate_geometry_inputs(dx)
    sc.credibility_tier = "engineering"
    with pytest.raises(SC.ScenarioError) as e:
        sc.validate_geometry_inputs(dx)
    assert "under-resolved" in e.value.message


@pytest.mark.parametrize("edit,needle", [
    (lambda r: r["impellers"][0].update(clearance_from_bottom_m=0.19), "below the liquid surface"),
    (lambda r: r["impellers"][0].update(diameter_m=0.25), "fit inside the vessel"),
    (lambda r: r["baffles"].append({"count": 3, "width_m": 0.02, "thickness_m": 0.004, "wall_attached": True}),
     "use count = 4"),
    (lambda r: r["baffles"].append({"count": 4, "width_m": 0.02, "thickness_m": 0.004, "wall_attached": False}),
     "detached baffle"),
    (lambda r: r["baffles"].append({"count": 4, "width_m": 0.09, "thickness_m": 0.02, "wall_attached": True}),
     "inside the tank radius"),
])
def test_geometry_refusals(edit, needle):
    d = base_scenario()
    edit(d["reactor"])
    sc = parse(d)
    with pytest.raises(SC.ScenarioError) as e:
        sc.validate_geometry_inputs(sc.unit_report_with_diagnostics()["lattice"]["dx_m"])
    assert needle in e.value.message


def test_stl_import_and_marine_refused_at_build():
    d = base_scenario()
    d["reactor"]["stl_import"] = {"path": "tank.stl"}
    sc = parse(d)  # parses, as in the reference
    with pytest.raises(SC.ScenarioError) as e:
        sc.build_solver()
    assert e.value.kind == "not_implemented" and "geometry-import" in e.value.message


def test_example_scenario_parses_and_maps():
    """examples/bioprocess/stirred_tank_screening.json: 48^3 Rushton tank,
    water at 250 rpm, passive scalar pulse, cell tracer."""
    sc = SC.BioprocessScenario.from_file(os.path.join(GOLDEN, "bioprocess_stirred_tank_screening.json"))
    rep = sc.compute_unit_report()
    lat, g = rep["lattice"], rep["groups"]
    assert_close(g["reynolds"], 1000.0 * (250 / 60) * 0.0455 ** 2 / 0.001)
    assert_close(lat["dx_m"], 0.1366 / 48)
    assert 0.5 < lat["tau_lu"] < 0.55  # water at this grid: near tau = 1/2
    assert "TAU_NEAR_HALF" in codes(rep, False)
    assert "STOKES_UNAVAILABLE" in codes(rep, False)  # cells present
    assert g["schmidt"] == pytest.approx(1e-6 / 1e-9)
    dx = lat["dx_m"]
    sc.validate_geometry_inputs(dx)
    assert math.isclose(sc.reactor["impellers"][0]["diameter_m"] / dx, 0.0455 / (0.1366 / 48))
