"""Bioprocess scenario ingestion: JSON -> unit mapping -> stirred-tank
geometry -> a ready-to-step ``Solver`` (the callers of the hot path).

Mirrors lbm-scenario's ``BioprocessScenario`` (crates/lbm-scenario/src/
bioprocess.rs:9-440), its unit report (units.rs:98-303, 340-521), the
stirred-tank geometry builder of lbm-core (geometry.rs:207-906) and the
single-phase runner setup (crates/lbm-cli/src/runner.rs:76-150, 955-1129,
1286-1425, 1514-1545).  Parsing keeps the reference's field names, its
``deny_unknown_fields`` strictness on the sections this path reads, and its
refusals with the same ``UnsupportedReason`` kinds.  Masks and markers are
built by the library's device / host geometry entry points; the solver is
the HIP one.  Nothing here computes a flow on the CPU.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np

# units.rs:8-33
TAU_NON_POSITIVE_VISCOSITY_THRESHOLD = 0.5
TAU_NEAR_HALF_WARN_THRESHOLD = 0.55
MACH_LATTICE_WARN_THRESHOLD = 0.10
MACH_LATTICE_REJECT_THRESHOLD = 0.30
CAHN_MIN_THRESHOLD = 0.02
CAHN_MAX_THRESHOLD = 0.5
BUBBLE_DIAMETER_MIN_CELLS = 4.0
BIOPROCESS_GRID_REYNOLDS_WARN_THRESHOLD = 15.0
GRID_SCHMIDT_WARN_THRESHOLD = 25.0
LATTICE_CS2 = 1.0 / 3.0
G_M_S2 = 9.80665
# geometry.rs:10-16
STIRRED_TANK_MIN_CELLS = 32.0
STIRRED_TANK_SCREENING_MIN_CELLS = 12.0
SPARGER_ORIFICE_MIN_CELLS = 3.0
RUSHTON_BLADE_THICKNESS_FRACTION = 0.04
PITCHED_BLADE_ANGLE_DEG = 45.0
PITCHED_BLADE_THICKNESS_FRACTION = 0.035


class ScenarioError(Exception):
    """``BioprocessScenarioError`` (bioprocess.rs:236-279): a message and an
    ``UnsupportedReason`` (solver.rs:172-180) as a dict with ``kind`` =
    not_implemented | out_of_validity_range | missing_dependency |
    evidence_gate_failed."""

    def __init__(self, message: str, reason: dict):
        super().__init__(f"{message} ({reason})")
        self.message = message
        self.reason = reason

    @property
    def kind(self) -> str:
        return self.reason["kind"]


def _range(message, detail):
    return ScenarioError(message, {"kind": "out_of_validity_range", "detail": detail})


def _missing(message, depends_on):
    return ScenarioError(message, {"kind": "missing_dependency", "depends_on": depends_on})


def _not_impl(message):
    return ScenarioError(message, {"kind": "not_implemented"})


def _gate(message, missing):
    return ScenarioError(message, {"kind": "evidence_gate_failed", "missing": list(missing)})


# ---------------------------------------------------------------------------
# schema (serde rename_all = snake_case, deny_unknown_fields)
# ---------------------------------------------------------------------------
_TOP = {"version", "name", "credibility_tier", "reactor", "fluids", "operation", "physics", "cells", "qoi", "run",
        "outputs"}
_TOP_REQUIRED = _TOP - {"cells"}
_TIERS = ("screening", "engineering", "evidence")
_REACTOR = {"kind", "vessel_diameter_m", "liquid_height_m", "working_volume_m3", "bottom", "stl_import", "impellers",
            "baffles", "spargers"}
_IMPELLER = {
    "rushton": {"kind", "diameter_m", "clearance_from_bottom_m", "rotational_speed_rpm", "blade_count"},
    "pitched_blade": {"kind", "diameter_m", "clearance_from_bottom_m", "rotational_speed_rpm", "blade_count"},
    "marine": {"kind", "diameter_m", "clearance_from_bottom_m", "rotational_speed_rpm", "blade_count"},
    "custom_marker_set": {"kind", "markers", "rotational_speed_rpm"},
}
_BAFFLE = {"count", "width_m", "thickness_m", "wall_attached"}
_SPARGER = {  # (required, optional): serde treats a missing Option field as None
    "ring": ({"kind", "center_z_m", "outer_radius_m", "orifice_count", "orifice_diameter_m", "inlet_phase"},
             {"raw_phi_boundary_fields", "gas_volumetric_flow_m3_per_s", "vvm"}),
    "pipe": ({"kind", "center_z_m", "length_m", "diameter_m", "inlet_phase"},
             {"axis", "orifice_count", "orifice_diameter_m", "raw_phi_boundary_fields",
              "gas_volumetric_flow_m3_per_s", "vvm"}),
    "point_orifices": ({"kind", "center_z_m", "positions", "inlet_phase"},
                       {"orifice_diameter_m", "raw_phi_boundary_fields", "gas_volumetric_flow_m3_per_s", "vvm"}),
}
_FLUIDS = {"liquid_density_kg_m3", "liquid_viscosity_pa_s", "gas_density_kg_m3", "gas_viscosity_pa_s",
           "surface_tension_n_m", "oxygen_diffusivity_m2_per_s", "henry_constant"}
_RUN = {"steps", "dt_s", "grid_nx", "grid_ny", "grid_nz", "backend", "precision", "lattice"}
_RUN_REQUIRED = {"steps", "dt_s", "grid_nx", "grid_ny", "grid_nz"}
_OUTPUTS = {"manifest_path", "fields_every_n_steps", "probes_every_n_steps", "emit_qoi_json", "emit_qoi_csv"}
_PHYSICS_KINDS = {"single_phase", "resolved_phase_field", "point_bubble", "hybrid", "passive_scalar", "oxygen",
                  "cell_tracer"}
_PULSE = {"center_m", "radius_m", "concentration"}


def _keys(obj, allowed, required, where):
    if not isinstance(obj, dict):
        raise _range(f"{where} must be an object", "invalid bioprocess scenario JSON")
    unknown = sorted(set(obj) - set(allowed))
    if unknown:
        raise _range(f"unknown field `{unknown[0]}` in {where}", "invalid bioprocess scenario JSON")
    miss = sorted(set(required) - set(obj))
    if miss:
        raise _range(f"missing field `{miss[0]}` in {where}", "invalid bioprocess scenario JSON")


def _num(v, where):
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise _range(f"{where} must be a number", "invalid bioprocess scenario JSON")
    return float(v)


def _opt_num(obj, key, where):
    v = obj.get(key)
    return None if v is None else _num(v, f"{where}.{key}")


def _uint(v, where):
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise _range(f"{where} must be a non-negative integer", "invalid bioprocess scenario JSON")
    return int(v)


@dataclass
class BioprocessScenario:
    """Parsed scenario; the raw sections are kept as dicts (``raw``)."""

    version: str
    name: str
    credibility_tier: str
    reactor: dict
    fluids: dict
    physics: list
    run: dict
    qoi: dict
    operation: dict
    outputs: dict
    cells: dict | None = None
    raw: dict = field(default_factory=dict, repr=False)

    # ---- parsing ----------------------------------------------------------
    @classmethod
    def from_json_str(cls, text: str) -> "BioprocessScenario":
        """``BioprocessScenario::from_json_str`` (bioprocess.rs:27-32):
        parse, then ``validate``."""
        try:
            doc = json.loads(text)
        except json.JSONDecodeError as e:
            raise _range(str(e), "invalid bioprocess scenario JSON") from None
        return cls.from_dict(doc)

    @classmethod
    def from_file(cls, path: str) -> "BioprocessScenario":
        with open(path) as fh:
            return cls.from_json_str(fh.read())

    @classmethod
    def from_dict(cls, doc: dict) -> "BioprocessScenario":
        _keys(doc, _TOP, _TOP_REQUIRED, "scenario")
        if doc["credibility_tier"] not in _TIERS:
            raise _range(f"unknown credibility_tier {doc['credibility_tier']!r}", "invalid bioprocess scenario JSON")
        reactor = _parse_reactor(doc["reactor"])
        _keys(doc["fluids"], _FLUIDS, {"liquid_density_kg_m3", "liquid_viscosity_pa_s"}, "fluids")
        fluids = {k: (None if doc["fluids"].get(k) is None else _num(doc["fluids"][k], f"fluids.{k}"))
                  for k in _FLUIDS}
        physics = doc["physics"]
        if isinstance(physics, dict):  # a single model is accepted untagged (bioprocess.rs PhysicsSpec)
            physics = [physics]
        if not isinstance(physics, list):
            raise _range("physics must be a list of models", "invalid bioprocess scenario JSON")
        for m in physics:
            if not isinstance(m, dict) or m.get("kind") not in _PHYSICS_KINDS:
                raise _range(f"unknown physics model {m!r}", "invalid bioprocess scenario JSON")
            if m["kind"] == "passive_scalar":
                _keys(m, {"kind", "diffusivity_m2_per_s", "initial_pulse"}, {"kind", "diffusivity_m2_per_s"},
                      "physics.passive_scalar")
                _num(m["diffusivity_m2_per_s"], "physics.passive_scalar.diffusivity_m2_per_s")
                if m.get("initial_pulse") is not None:
                    _keys(m["initial_pulse"], _PULSE, _PULSE, "physics.passive_scalar.initial_pulse")
            if m["kind"] == "single_phase":
                _keys(m, {"kind"}, {"kind"}, "physics.single_phase")
        _keys(doc["run"], _RUN, _RUN_REQUIRED, "run")
        run = dict(doc["run"])
        run["steps"] = _uint(run["steps"], "run.steps")
        run["dt_s"] = _num(run["dt_s"], "run.dt_s")
        for k in ("grid_nx", "grid_ny", "grid_nz"):
            run[k] = _uint(run[k], f"run.{k}")
        for k, allowed in (("backend", ("auto", "cpu", "gpu")), ("precision", ("f32", "f64")),
                           ("lattice", ("d2q9", "d3q19", "d3q27"))):
            if run.get(k) is not None and run[k] not in allowed:
                raise _range(f"unknown run.{k} {run[k]!r}", "invalid bioprocess scenario JSON")
        _keys(doc["outputs"], _OUTPUTS, {"manifest_path", "emit_qoi_json", "emit_qoi_csv"}, "outputs")
        sc = cls(version=doc["version"], name=doc["name"], credibility_tier=doc["credibility_tier"],
                 reactor=reactor, fluids=fluids, physics=physics, run=run, qoi=doc["qoi"] or {},
                 operation=doc["operation"], outputs=doc["outputs"], cells=doc.get("cells"), raw=doc)
        sc.validate()
        return sc

    # ---- model predicates (bioprocess.rs:879-900, units.rs:460-473) ----------
    def _kinds(self):
        return [m["kind"] for m in self.physics]

    def has_oxygen(self) -> bool:
        return "oxygen" in self._kinds()

    def has_gas_model(self) -> bool:
        return any(k in ("resolved_phase_field", "point_bubble", "hybrid") for k in self._kinds())

    def has_scalar_model(self) -> bool:
        return any(k in ("passive_scalar", "oxygen") for k in self._kinds())

    def has_particle_model(self) -> bool:
        return self.cells is not None or "cell_tracer" in self._kinds()

    def primary_phase_field(self):
        for m in self.physics:
            if m["kind"] == "resolved_phase_field":
                return float(m["interface_width_m"]), float(m["mobility_m2_per_s"])
            if m["kind"] == "hybrid":
                pf = m["phase_field"]
                return float(pf["interface_width_m"]), float(pf["mobility_m2_per_s"])
        return None

    def primary_scalar_diffusivity(self):
        for m in self.physics:
            if m["kind"] == "passive_scalar":
                return float(m["diffusivity_m2_per_s"])
        if self.has_oxygen():
            return self.fluids.get("oxygen_diffusivity_m2_per_s")
        return None

    # ---- validation (bioprocess.rs:34-133, 288-337) -------------------------
    def validate(self):
        qoi = self.qoi
        if qoi.get("kla") is not None and not self.has_oxygen():
            raise _missing("kLa QOI requires oxygen physics", "oxygen")
        if self.has_oxygen():
            for k in ("oxygen_diffusivity_m2_per_s", "henry_constant"):
                v = self.fluids.get(k)
                if v is None or not (math.isfinite(v) and v > 0.0):
                    raise _range(f"oxygen physics requires fluids.{k} to be finite and > 0",
                                 f"fluids.{k} must be provided for oxygen physics")
        if self.reactor["spargers"] and not self.has_gas_model():
            raise _missing("spargers require a gas model in physics", "gas_model_in_physics")
        cal, hold = qoi.get("calibration_dataset_id"), qoi.get("holdout_dataset_id")
        if cal is not None and hold is not None and cal == hold:
            raise _range("calibration and holdout datasets must differ", "dataset reuse: all_requested_qois")
        if self.credibility_tier == "evidence" and not (cal is not None and hold is not None):
            raise _gate("evidence tier requires calibration and holdout dataset registry references",
                        ["calibration_and_holdout_dataset_registry"])
        for s in self.reactor["spargers"]:
            if s.get("inlet_phase") != "gas":
                raise _range(f"sparger inlet phase must be gas, got {s.get('inlet_phase')}",
                             "sparger inlet_phase must be gas")
        if self.run.get("lattice") == "d2q9":
            raise _range("bioprocess stirred-tank scenarios require a 3D lattice",
                         "run.lattice=d2q9 is outside the 3D stirred-tank validity range")
        if self.reactor["bottom"] == "dished":
            raise _not_impl("dished-bottom stirred tanks are not implemented for M0")
        if min(self.run["grid_nx"], self.run["grid_ny"], self.run["grid_nz"]) < 3:
            raise _range("stirred-tank grid dimensions must preserve a 1-cell solid rim",
                         "run.grid_nx, run.grid_ny and run.grid_nz must each be >= 3")
        for s in self.reactor["spargers"]:
            if s.get("raw_phi_boundary_fields"):
                raise _range("raw phi boundary fields are not accepted for sparger geometry",
                             "spargers declare geometry and gas metadata only; BCFD-046 owns gas injection")
        for m in self.physics:
            if m["kind"] in ("resolved_phase_field",) and m.get("dynamic_contact_angle") is not None:
                raise _not_impl("dynamic contact angle is not implemented")
            if m["kind"] == "hybrid" and m["phase_field"].get("dynamic_contact_angle") is not None:
                raise _not_impl("dynamic contact angle is not implemented")

    # ---- unit mapping (units.rs:98-303) ---------------------------------------
    def unit_report_with_diagnostics(self) -> dict:
        """``bioprocess_unit_report_unchecked``: the mapping and every
        feasibility issue, without refusing."""
        r = self.reactor
        if not r["impellers"]:
            raise _missing("at least one impeller is required for bioprocess unit mapping", "reactor.impellers")
        _require_positive("reactor.vessel_diameter_m", r["vessel_diameter_m"])
        _require_positive("reactor.liquid_height_m", r["liquid_height_m"])
        imp = r["impellers"][0]
        if imp["kind"] == "custom_marker_set":
            raise _missing("unit mapping requires a parametric impeller diameter", "impeller.diameter_m")
        D = imp["diameter_m"]
        _require_positive("impeller.diameter_m", D)
        _require_positive("impeller.rotational_speed_rpm", imp["rotational_speed_rpm"])
        fl, run = self.fluids, self.run
        _require_positive("fluids.liquid_density_kg_m3", fl["liquid_density_kg_m3"])
        _require_positive("fluids.liquid_viscosity_pa_s", fl["liquid_viscosity_pa_s"])
        _require_positive("run.dt_s", run["dt_s"])
        if run["grid_nx"] == 0 or run["grid_ny"] == 0 or run["grid_nz"] == 0:
            raise _range("run grid dimensions must be positive",
                         "run.grid_nx, run.grid_ny and run.grid_nz must all be > 0")
        rho, mu, dt = fl["liquid_density_kg_m3"], fl["liquid_viscosity_pa_s"], run["dt_s"]
        n_hz = imp["rotational_speed_rpm"] / 60.0
        u_ref = n_hz * D
        nu = mu / rho
        dx = r["vessel_diameter_m"] / float(run["grid_nx"])
        u_ref_lu = u_ref * dt / dx
        nu_lu = nu * dt / (dx * dx)
        tau_lu = 3.0 * nu_lu + 0.5
        mach = u_ref_lu / math.sqrt(LATTICE_CS2)
        reynolds = rho * n_hz * D ** 2 / mu
        froude = n_hz ** 2 * D / G_M_S2
        sigma, rho_g = fl.get("surface_tension_n_m"), fl.get("gas_density_kg_m3")
        drho = None if rho_g is None else rho - rho_g
        weber = None if sigma is None else rho * n_hz ** 2 * D ** 3 / sigma
        eotvos = None if (sigma is None or drho is None) else G_M_S2 * drho * D ** 2 / sigma
        morton = None if (sigma is None or drho is None) else G_M_S2 * mu ** 4 * drho / (rho ** 2 * sigma ** 3)
        dsc = self.primary_scalar_diffusivity()
        schmidt = None if dsc is None else nu / dsc
        peclet = None if dsc is None else u_ref * r["vessel_diameter_m"] / dsc
        dsc_lu = None if dsc is None else dsc * dt / (dx * dx)
        phase = self.primary_phase_field()
        cahn = None if phase is None else phase[0] / r["vessel_diameter_m"]
        peclet_phi = None if phase is None else u_ref * phase[0] / phase[1]

        warnings, rejections = [], []
        if mach > MACH_LATTICE_REJECT_THRESHOLD:
            rejections.append(_issue("MA_LATTICE_TOO_HIGH", "lattice Mach exceeds the hard low-Mach cap", mach,
                                     MACH_LATTICE_REJECT_THRESHOLD))
        elif mach > MACH_LATTICE_WARN_THRESHOLD:
            warnings.append(_issue("MA_LATTICE_HIGH",
                                   "lattice Mach is high enough for compressibility error to matter", mach,
                                   MACH_LATTICE_WARN_THRESHOLD))
        if tau_lu <= TAU_NON_POSITIVE_VISCOSITY_THRESHOLD:
            rejections.append(_issue("TAU_NON_POSITIVE_VISCOSITY",
                                     "BGK relaxation time implies non-positive lattice viscosity", tau_lu,
                                     TAU_NON_POSITIVE_VISCOSITY_THRESHOLD))
        elif tau_lu <= TAU_NEAR_HALF_WARN_THRESHOLD:
            warnings.append(_issue("TAU_NEAR_HALF", "BGK relaxation time is numerically marginal near tau = 0.5",
                                   tau_lu, TAU_NEAR_HALF_WARN_THRESHOLD))
        if cahn is not None:
            if cahn > CAHN_MAX_THRESHOLD:
                rejections.append(_issue("INTERFACE_WIDTH_TOO_LARGE",
                                         "phase-field interface width is too large relative to the vessel "
                                         "diameter", cahn, CAHN_MAX_THRESHOLD))
            elif cahn < CAHN_MIN_THRESHOLD:
                rejections.append(_issue("INTERFACE_WIDTH_TOO_THIN",
                                         "phase-field interface width is too thin relative to the vessel diameter",
                                         cahn, CAHN_MIN_THRESHOLD))
        if phase is not None:
            for d in _sparger_diameters(r["spargers"]):
                cells = d / dx
                if cells < BUBBLE_DIAMETER_MIN_CELLS:
                    rejections.append(_issue("BUBBLE_UNDER_RESOLVED",
                                             "resolved gas injection diameter spans too few lattice cells", cells,
                                             BUBBLE_DIAMETER_MIN_CELLS))
        grid_re = u_ref_lu / nu_lu
        if grid_re > BIOPROCESS_GRID_REYNOLDS_WARN_THRESHOLD:
            warnings.append(_issue("GRID_REYNOLDS_HIGH", "cell-scale Reynolds number is high for the selected "
                                   "mapping", grid_re, BIOPROCESS_GRID_REYNOLDS_WARN_THRESHOLD))
        if dsc_lu is not None and nu_lu / dsc_lu > GRID_SCHMIDT_WARN_THRESHOLD:
            warnings.append(_issue("SCALAR_DIFFUSION_UNSTABLE", "lattice scalar diffusion is marginal for the "
                                   "selected mapping", nu_lu / dsc_lu, GRID_SCHMIDT_WARN_THRESHOLD))
        # add_optional_group_reasons, units.rs:403-458
        if self.has_gas_model() and weber is None:
            warnings.append(_issue("WEBER_UNAVAILABLE",
                                   "Weber number not emitted because surface_tension_n_m is absent"))
        if self.has_gas_model() and eotvos is None:
            warnings.append(_issue("EOTVOS_UNAVAILABLE", "Eotvos number not emitted because surface_tension_n_m "
                                   "or gas_density_kg_m3 is absent"))
        if self.has_gas_model() and morton is None:
            warnings.append(_issue("MORTON_UNAVAILABLE", "Morton number not emitted because surface_tension_n_m "
                                   "or gas_density_kg_m3 is absent"))
        if self.has_scalar_model() and dsc is None:
            warnings.append(_issue("SCALAR_GROUPS_UNAVAILABLE", "Schmidt and Peclet numbers not emitted because "
                                   "scalar diffusivity is absent"))
        if self.has_particle_model():
            warnings.append(_issue("STOKES_UNAVAILABLE", "Stokes number not emitted because particle species "
                                   "density and diameter are absent from the schema"))
        return {
            "lattice": {"grid_nx": run["grid_nx"], "grid_ny": run["grid_ny"], "grid_nz": run["grid_nz"],
                        "dx_m": dx, "dt_s": dt, "u_ref_lu": u_ref_lu, "nu_lu": nu_lu, "tau_lu": tau_lu,
                        "cs2_lu": LATTICE_CS2},
            "groups": {"reynolds": reynolds, "froude": froude, "weber": weber, "eotvos": eotvos, "morton": morton,
                       "schmidt": schmidt, "peclet": peclet, "stokes": None, "mach_lattice": mach, "cahn": cahn,
                       "peclet_phi": peclet_phi},
            "feasibility": {"warnings": warnings, "rejections": rejections},
            "matching_priority": self._matching_priority(weber, eotvos, dsc),
        }

    def compute_unit_report(self) -> dict:
        """``bioprocess_unit_report`` (units.rs:98-115): refuses a mapping
        with any feasibility rejection."""
        rep = self.unit_report_with_diagnostics()
        rej = rep["feasibility"]["rejections"]
        if rej:
            detail = "; ".join(f"{i['code']}: {i['message']}" for i in rej)
            raise _range("bioprocess unit feasibility rejected the scenario", detail)
        return rep

    def _matching_priority(self, weber, eotvos, dsc):  # units.rs:475-521
        relaxed, missing = [], []
        if self.primary_phase_field() is not None:
            if self.fluids.get("gas_density_kg_m3") is not None and weber is not None and eotvos is not None:
                relaxed.append("rho_gas/rho_liquid and We/Eo fall out of the Re-matched lattice mapping")
            else:
                missing.append("rho_gas/rho_liquid and We/Eo not enforced because gas density or surface "
                               "tension is absent")
        else:
            missing.append("rho_gas/rho_liquid and We/Eo not enforced because phase field is off")
        relaxed.append("Fr falls out of the Re-matched lattice mapping when free-surface behavior is relevant")
        if self.has_scalar_model():
            if dsc is not None:
                relaxed.append("Sc/Pe/Da fall out of the Re-matched lattice mapping for active scalars")
            else:
                missing.append("Sc/Pe/Da not enforced because scalar diffusivity is absent")
        else:
            missing.append("Sc/Pe/Da not enforced because no scalar or reaction model is active")
        if self.has_particle_model():
            missing.append("St not enforced because particle species density and diameter are absent")
        else:
            missing.append("St not enforced because no particle model is active")
        return {"achieved": ["Re"], "relaxed": relaxed, "missing": missing}

    # ---- geometry (geometry.rs:224-268, bioprocess.rs:143-153) ---------------
    def resolution_floor(self) -> float:
        """stirred_tank_resolution_floor, bioprocess.rs:452-459."""
        return STIRRED_TANK_SCREENING_MIN_CELLS if self.credibility_tier == "screening" else STIRRED_TANK_MIN_CELLS

    def validate_geometry_inputs(self, dx_m: float):
        """The checks of build_stirred_tank_geometry_with_min_cells before any
        mask is written (geometry.rs:224-244, 368-570)."""
        r, run = self.reactor, self.run
        dims = (run["grid_nx"], run["grid_ny"], run["grid_nz"])
        if min(dims) < 3:
            raise _range("stirred-tank grid dimensions must preserve a 1-cell solid rim",
                         "run.grid_nx, run.grid_ny and run.grid_nz must each be >= 3")
        if not (math.isfinite(dx_m) and dx_m > 0.0):
            raise _range("grid spacing must be finite and positive", "dx_m must be finite and > 0")
        T, H = r["vessel_diameter_m"], r["liquid_height_m"]
        if not (math.isfinite(T) and T > 0.0):
            raise _range("vessel diameter must be finite and positive", "reactor.vessel_diameter_m must be > 0")
        if not (math.isfinite(H) and H > 0.0):
            raise _range("liquid height must be finite and positive", "reactor.liquid_height_m must be > 0")
        if r["bottom"] == "dished":
            raise _not_impl("dished-bottom stirred tanks are not implemented for M0")
        floor = self.resolution_floor()
        lmin = T
        for imp in r["impellers"]:
            if imp["kind"] != "custom_marker_set":
                lmin = min(lmin, imp["diameter_m"])
        cells = lmin / dx_m
        if cells < floor:
            raise _range("stirred-tank geometry is under-resolved",
                         f"min(impeller diameter, vessel diameter) / dx must be >= {floor}, got {cells:.3f}")
        for imp in r["impellers"]:
            if imp["kind"] == "custom_marker_set":
                _require_positive("impeller.rotational_speed_rpm", imp["rotational_speed_rpm"], geometry=True)
                _validate_custom_markers(T, H, imp["markers"])
                continue
            _require_positive("impeller.diameter_m", imp["diameter_m"], geometry=True)
            _require_positive("impeller.rotational_speed_rpm", imp["rotational_speed_rpm"], geometry=True)
            if imp["blade_count"] == 0:
                raise _range("impeller blade_count must be positive", "impeller.blade_count must be > 0")
            rad = 0.5 * imp["diameter_m"]
            c = imp["clearance_from_bottom_m"]
            if not (c > 0.0 and c + rad < H):
                raise _range("impeller must be above the tank bottom and below the liquid surface",
                             "impeller clearance and radius must fit inside the liquid volume")
            if imp["diameter_m"] >= T:
                raise _range("impeller diameter must fit inside the vessel",
                             "impeller.diameter_m must be less than reactor.vessel_diameter_m")
        for b in r["baffles"]:
            if b["count"] not in (0, 4):
                raise _not_impl("custom baffle placement is not implemented for M0; use count = 4")
            if b["count"] == 0:
                continue
            _require_positive("baffle.width_m", b["width_m"], geometry=True)
            _require_positive("baffle.thickness_m", b["thickness_m"], geometry=True)
            if not b["wall_attached"]:
                raise _not_impl("detached baffle placement is not implemented for M0")
            if b["width_m"] + b["thickness_m"] >= 0.5 * T:
                raise _range("baffles must fit inside the tank radius",
                             "baffle width plus thickness must be less than the tank radius")
        for s in r["spargers"]:
            _validate_sparger(T, H, dx_m, s)

    def build_geometry(self, ctx=None) -> "StirredTankGeometry":
        """``BioprocessScenario::build_geometry`` (bioprocess.rs:143-153 over
        geometry.rs:224-268): tank, baffle and sparger masks, impeller
        marker sets and their wall velocity stamp.  Masks come from the
        library's device kernel, markers from its host generator."""
        from . import solver as S

        rep = self.unit_report_with_diagnostics()
        dx = rep["lattice"]["dx_m"]
        self.validate_geometry_inputs(dx)
        r, run = self.reactor, self.run
        dims = (run["grid_nx"], run["grid_ny"], run["grid_nz"])
        T, H = r["vessel_diameter_m"], r["liquid_height_m"]
        n = dims[0] * dims[1] * dims[2]
        Q = 27 if run.get("lattice") == "d3q27" else 19
        probe = S.Solver(Q, dims, nu=0.1, ctx=ctx)
        try:
            probe.build_stirred_tank(dx, T, H, 0, 0.0, 0.0)
            tank_only = probe.solid().astype(bool)
            solid = tank_only.copy()
            for b in r["baffles"]:
                if b["count"] == 4:
                    probe.build_stirred_tank(dx, T, H, 4, b["width_m"], b["thickness_m"])
                    solid |= probe.solid().astype(bool)
        finally:
            probe.close()
        geo = StirredTankGeometry(dims=dims, dx_m=dx, solid=solid.astype(np.uint8),
                                  wall_velocity=np.zeros((n, 3)), baffle_mask=(solid & ~tank_only).astype(np.uint8),
                                  sparger_mask=np.zeros(n, np.uint8))
        for s in r["spargers"]:
            _stamp_sparger(dims, dx, H, s, geo)
        for imp in r["impellers"]:
            kind = imp["kind"]
            if kind == "marine":
                raise _not_impl("marine impeller marker generation is not implemented")
            if kind == "custom_marker_set":
                raise _not_impl("custom impeller marker-set generation is not implemented")
            angle, thick = ((0.0, RUSHTON_BLADE_THICKNESS_FRACTION) if kind == "rushton"
                            else (PITCHED_BLADE_ANGLE_DEG, PITCHED_BLADE_THICKNESS_FRACTION))
            omega = imp["rotational_speed_rpm"] * 2.0 * math.pi / 60.0  # rpm_to_rad_s
            pos, w, vel, c = S.radial_impeller(dims, dx, T, H, imp["diameter_m"], imp["clearance_from_bottom_m"],
                                               imp["blade_count"], angle, omega, thick)
            ms = ImpellerMarkerSet(kind=kind, center=c, omega=np.array([0.0, 0.0, omega]), positions=pos,
                                   weights=w, wall_velocity=vel)
            # stamp_impeller_wall_velocity, geometry.rs:908-920
            for p, v in zip(pos, vel):
                ix = [int(min(max(round(p[a]), 0.0), dims[a] - 1)) for a in range(3)]
                geo.wall_velocity[(ix[2] * dims[1] + ix[1]) * dims[0] + ix[0]] = v
            geo.impellers.append(ms)
        return geo

    # ---- runner setup (runner.rs:955-1014, 1286-1298, 1343-1378) ------------
    def build_solver(self, ctx=None, storage: str = "f32"):
        """The single-phase runner's solver: TRT with magic 3/16 at
        nu_lu, closed faces, the geometry's solid mask with wall velocity
        and impeller omega scaled to lattice time, the rotating IBM body and
        the passive scalar with its top-hat pulse.  Returns a
        ``ScenarioRun``."""
        from . import solver as S

        if self.reactor.get("stl_import") is not None:  # import_stl_geometry, bioprocess.rs:155-163, 370-380
            raise _not_impl("STL geometry import requires the geometry-import feature")
        if "single_phase" not in self._kinds():
            raise _range("bioprocess runner currently supports physics.kind single_phase",
                         "physics must include single_phase")
        rep = self.compute_unit_report()
        geo = self.build_geometry(ctx)
        if len(geo.impellers) != 1:
            raise _not_impl("this backend drives one impeller marker set per tank")
        dt = self.run["dt_s"]
        Q = 27 if self.run.get("lattice") == "d3q27" else 19
        wall_u = (geo.wall_velocity * dt).astype(np.float32)
        s = S.Solver(Q, geo.dims, nu=rep["lattice"]["nu_lu"], collision="trt", storage=storage,
                     solid=geo.solid, wall_u=wall_u, ctx=ctx)
        imp = geo.impellers[0]
        s.set_rotating_body(imp.positions, imp.weights, imp.center, tuple(imp.omega * dt))
        scalar = None
        for m in self.physics:
            if m["kind"] == "passive_scalar":
                d_lu = float(m["diffusivity_m2_per_s"]) * dt / rep["lattice"]["dx_m"] ** 2
                k = s.enable_scalar_distribution(d_lu, "hydro")
                pulse = m.get("initial_pulse")
                if pulse is not None:
                    s.set_scalar_concentration(k, top_hat_pulse(geo.dims, rep["lattice"]["dx_m"], geo.solid, pulse))
                scalar = k
                break
        return ScenarioRun(scenario=self, report=rep, geometry=geo, solver=s, scalar=scalar)


@dataclass
class ImpellerMarkerSet:
    """``ImpellerMarkerSet`` (geometry.rs:193-205): positions in lattice
    units, omega in rad/s, wall velocity in lattice lengths per second."""

    kind: str
    center: np.ndarray
    omega: np.ndarray
    positions: np.ndarray
    weights: np.ndarray
    wall_velocity: np.ndarray


@dataclass
class StirredTankGeometry:
    """``StirredTankGeometry`` (geometry.rs:167-191), compact cell order."""

    dims: tuple
    dx_m: float
    solid: np.ndarray
    wall_velocity: np.ndarray
    baffle_mask: np.ndarray
    sparger_mask: np.ndarray
    sparger_orifice_centers: list = field(default_factory=list)
    impellers: list = field(default_factory=list)

    def solid_count(self) -> int:
        return int(self.solid.sum())

    def fluid_count(self) -> int:
        return int(self.solid.size - self.solid.sum())


@dataclass
class ScenarioRun:
    """A built scenario: the HIP solver plus what the runner reports with it."""

    scenario: BioprocessScenario
    report: dict
    geometry: StirredTankGeometry
    solver: object
    scalar: int | None

    def run(self, steps: int | None = None, probe_every: int | None = None, graph: bool = True) -> dict:
        """The runner's step loop (runner.rs:1079-1129) over ``steps``
        (default run.steps): torque of every step in lattice and SI units
        (runner.rs:1400-1425), the scalar CV at every probe step, and the
        power QOIs over the last half of the run (runner.rs:1514-1545)."""
        from . import solver as S

        sc = self.scenario
        steps = sc.run["steps"] if steps is None else int(steps)
        probe = int(probe_every or sc.outputs.get("probes_every_n_steps") or 1)
        lat = self.report["lattice"]
        rho = sc.fluids["liquid_density_kg_m3"]
        dx, dt = lat["dx_m"], lat["dt_s"]
        torque_scale = rho * dx ** 5 / dt ** 2
        tq_lu = np.zeros(steps)
        cv = []
        done = 0
        while done < steps:
            chunk = min(probe - (done % probe), steps - done)
            tq, diag = self.solver.run_with_ibm(chunk, graph=graph)
            tq_lu[done:done + chunk] = -tq[:, 2]
            done += chunk
            if self.scalar is not None and (done % probe == 0 or done == steps):
                value = self.solver.scalar_cv(self.scalar)
                if value is not None:
                    cv.append((done * dt, value))
        rec_steps = np.array([k for k in range(1, steps + 1) if k % probe == 0 or k == steps])
        rec_tq = tq_lu[rec_steps - 1] * torque_scale
        start = sc.run["steps"] // 2 if steps == sc.run["steps"] else steps // 2
        win = rec_tq[rec_steps >= start]
        torque = float((win if win.size else rec_tq).mean()) if rec_tq.size else 0.0
        imp = sc.reactor["impellers"][0]
        omega = imp["rotational_speed_rpm"] * 2.0 * math.pi / 60.0
        power = S.power_qois(torque, omega, rho, imp["diameter_m"], sc.reactor["working_volume_m3"])
        out = {"steps": steps, "torque_lu": tq_lu, "torque_n_m": tq_lu * torque_scale, "record_steps": rec_steps,
               "mean_torque_n_m": torque, "power": power, "scalar_cv": cv}
        if cv:
            t = np.array([c[0] for c in cv])
            v = np.array([c[1] for c in cv])
            out["mixing_time"] = S.mixing_time_from_cv(t, v)
        return out

    def close(self):
        self.solver.close()


def top_hat_pulse(dims, dx, solid, pulse) -> np.ndarray:
    """The passive scalar's initial top-hat sphere (runner.rs:1343-1371):
    fluid cells whose centre lies within radius_m of center_m hold
    ``concentration``."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    px, py, pz = ((x.reshape(-1) + 0.5) * dx, (y.reshape(-1) + 0.5) * dx, (z.reshape(-1) + 0.5) * dx)
    c = pulse["center_m"]
    r = np.sqrt((px - c[0]) ** 2 + (py - c[1]) ** 2 + (pz - c[2]) ** 2)
    v = np.where((r <= pulse["radius_m"]) & (np.asarray(solid).reshape(-1) == 0), float(pulse["concentration"]), 0.0)
    return v


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _issue(code, message, value=None, threshold=None):
    return {"code": code, "message": message, "value": value, "threshold": threshold}


def _require_positive(name, value, geometry=False):
    if value is None or not (math.isfinite(value) and value > 0.0):
        detail = f"{name} must be finite and > 0" if geometry else f"{name} must be finite and positive"
        raise _range(f"{name} must be finite and positive", detail)


def _parse_reactor(r):
    if not isinstance(r, dict) or r.get("kind") != "stirred_tank":
        raise _range("reactor.kind must be stirred_tank", "invalid bioprocess scenario JSON")
    _keys(r, _REACTOR, {"kind", "vessel_diameter_m", "liquid_height_m", "working_volume_m3", "impellers", "baffles",
                        "spargers"}, "reactor")
    out = dict(r)
    for k in ("vessel_diameter_m", "liquid_height_m", "working_volume_m3"):
        out[k] = _num(r[k], f"reactor.{k}")
    out["bottom"] = r.get("bottom") or "flat"
    if out["bottom"] not in ("flat", "dished"):
        raise _range(f"unknown reactor.bottom {out['bottom']!r}", "invalid bioprocess scenario JSON")
    if r.get("stl_import") is not None:
        _keys(r["stl_import"], {"path", "labels_path"}, {"path"}, "reactor.stl_import")
    imps = []
    for imp in r["impellers"]:
        kind = imp.get("kind") if isinstance(imp, dict) else None
        if kind not in _IMPELLER:
            raise _range(f"unknown impeller kind {kind!r}", "invalid bioprocess scenario JSON")
        _keys(imp, _IMPELLER[kind], _IMPELLER[kind], f"impeller.{kind}")
        d = dict(imp)
        d["rotational_speed_rpm"] = _num(imp["rotational_speed_rpm"], "impeller.rotational_speed_rpm")
        if kind != "custom_marker_set":
            d["diameter_m"] = _num(imp["diameter_m"], "impeller.diameter_m")
            d["clearance_from_bottom_m"] = _num(imp["clearance_from_bottom_m"], "impeller.clearance_from_bottom_m")
            d["blade_count"] = _uint(imp["blade_count"], "impeller.blade_count")
        imps.append(d)
    out["impellers"] = imps
    baffles = []
    for b in r["baffles"]:
        _keys(b, _BAFFLE, _BAFFLE, "baffle")
        baffles.append({"count": _uint(b["count"], "baffle.count"), "width_m": _num(b["width_m"], "baffle.width_m"),
                        "thickness_m": _num(b["thickness_m"], "baffle.thickness_m"),
                        "wall_attached": bool(b["wall_attached"])})
    out["baffles"] = baffles
    sps = []
    for s in r["spargers"]:
        kind = s.get("kind") if isinstance(s, dict) else None
        if kind not in _SPARGER:
            raise _range(f"unknown sparger kind {kind!r}", "invalid bioprocess scenario JSON")
        req, opt = _SPARGER[kind]
        _keys(s, req | opt, req, f"sparger.{kind}")
        sps.append(dict(s))
    out["spargers"] = sps
    return out


def _sparger_diameters(spargers):  # units.rs:384-401
    out = []
    for s in spargers:
        if s["kind"] == "ring":
            out.append(float(s["orifice_diameter_m"]))
        elif s["kind"] == "pipe":
            out.append(float(s["orifice_diameter_m"] if s.get("orifice_diameter_m") is not None else s["diameter_m"]))
        elif s.get("orifice_diameter_m") is not None:
            out.append(float(s["orifice_diameter_m"]))
    return out


def _sparger_template(s):
    """SpargerSpec -> SpargerTemplate (bioprocess.rs From impl): pipe axis
    defaults to x, orifice count to 1 and orifice diameter to the pipe's."""
    gas = s.get("gas_volumetric_flow_m3_per_s")
    if s["kind"] == "ring":
        return dict(center_z_m=s["center_z_m"], outer_radius_m=s["outer_radius_m"], orifice_count=s["orifice_count"],
                    orifice_diameter_m=s["orifice_diameter_m"], gas=gas, centers=None)
    if s["kind"] == "pipe":
        return dict(center_z_m=s["center_z_m"], length_m=s["length_m"], diameter_m=s["diameter_m"],
                    axis=s.get("axis") or "x", orifice_count=s.get("orifice_count") or 1,
                    orifice_diameter_m=s.get("orifice_diameter_m") or s["diameter_m"], gas=gas)
    return dict(center_z_m=s["center_z_m"], positions=s["positions"],
                orifice_diameter_m=s.get("orifice_diameter_m") or 0.0, gas=gas, orifice_count=len(s["positions"]))


def _validate_sparger(T, H, dx, s):  # geometry.rs:510-624
    t = _sparger_template(s)
    cz = float(t["center_z_m"])
    if not (math.isfinite(cz) and 0.0 < cz < H):
        raise _range("sparger centre must be below the liquid surface and above the tank bottom",
                     "sparger center_z_m must lie inside the liquid volume")
    if t["gas"] is None or not (math.isfinite(t["gas"]) and t["gas"] > 0.0):
        raise _range("sparger gas volumetric flow must be positive",
                     "gas_volumetric_flow_m3_per_s must be finite and > 0")
    if t["orifice_count"] == 0:
        raise _range("sparger orifice count must be positive", "orifice_count must be > 0")
    _require_positive("sparger.orifice_diameter_m", t["orifice_diameter_m"], geometry=True)
    cells = t["orifice_diameter_m"] / dx
    if cells < SPARGER_ORIFICE_MIN_CELLS:
        raise _range("orifice under-resolved for resolved injection (BCFD-046)",
                     f"orifice_diameter_m / dx must be >= {SPARGER_ORIFICE_MIN_CELLS}, got {cells:.3f}")
    rad = 0.5 * T
    if s["kind"] == "ring":
        _require_positive("sparger.outer_radius_m", t["outer_radius_m"], geometry=True)
        if t["outer_radius_m"] >= rad:
            raise _range("ring sparger must be inside the tank", "outer_radius_m must be less than the tank radius")
    elif s["kind"] == "pipe":
        _require_positive("sparger.length_m", t["length_m"], geometry=True)
        _require_positive("sparger.diameter_m", t["diameter_m"], geometry=True)
        if 0.5 * t["length_m"] + 0.5 * t["diameter_m"] >= rad:
            raise _range("pipe sparger must be inside the tank",
                         "length_m/2 plus diameter_m/2 must be less than the tank radius")
    else:
        for p in t["positions"]:
            if not all(math.isfinite(v) for v in p):
                raise _range("point sparger positions must be finite", "positions entries must be finite")
            if not (math.hypot(p[0], p[1]) < rad and 0.0 < p[2] < H):
                raise _range("point sparger positions must lie inside the tank",
                             "each point-orifice position must be inside the liquid volume")


def _stamp_sparger(dims, dx, H, s, geo):
    """apply_sparger_mask / stamp_orifice (geometry.rs:720-812): the cells
    whose centre lies within half an orifice diameter of each orifice
    centre (tank-centred x, y; absolute z) are marked; the centres are
    recorded."""
    t = _sparger_template(s)
    centers = []
    if s["kind"] == "ring":
        m = t["orifice_count"]
        for k in range(m):
            th = 2.0 * math.pi * k / m
            centers.append((t["outer_radius_m"] * math.cos(th), t["outer_radius_m"] * math.sin(th), t["center_z_m"]))
    elif s["kind"] == "pipe":
        m = t["orifice_count"]
        denom = max(m, 2) - 1
        for k in range(m):
            sk = 0.0 if m == 1 else -0.5 * t["length_m"] + t["length_m"] * k / denom
            centers.append((sk, 0.0, t["center_z_m"]) if t["axis"] == "x" else (0.0, sk, t["center_z_m"]))
    else:
        centers = [tuple(p) for p in t["positions"]]
    nx, ny, nz = dims
    cx, cy = 0.5 * nx * dx, 0.5 * ny * dx
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
    px, py, pz = (x + 0.5) * dx, (y + 0.5) * dx, (z + 0.5) * dx
    inner = (x >= 1) & (x <= nx - 2) & (y >= 1) & (y <= ny - 2) & (z >= 1) & (z <= nz - 2) & (pz < H)
    rad = 0.5 * t["orifice_diameter_m"]
    for c in centers:
        cc = (cx + c[0], cy + c[1], c[2])
        geo.sparger_orifice_centers.append(cc)
        d = np.sqrt((px - cc[0]) ** 2 + (py - cc[1]) ** 2 + (pz - cc[2]) ** 2)
        geo.sparger_mask[inner & (d <= rad)] = 1


def _validate_custom_markers(T, H, markers):  # geometry.rs:451-477
    if not markers:
        raise _range("custom impeller marker set must not be empty",
                     "impeller.markers must contain at least one marker")
    for m in markers:
        if not all(math.isfinite(v) for v in m):
            raise _range("custom impeller markers must be finite", "impeller.markers entries must be finite")
        if not (math.hypot(m[0], m[1]) < 0.5 * T and 0.0 < m[2] < H):
            raise _range("custom impeller marker must lie inside the liquid volume",
                         "custom marker radius and z must fit inside the tank liquid")
