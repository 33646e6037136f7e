#!/usr/bin/env python3
"""Benchmark: MLUPS and HBM-roofline fraction of the D3Q27 central-moment
stirred tank at 512^3 on one MI355X.

One step is the reference runner's step (runner.rs:1079-1129): clear the
body-force field, rotating-IBM direct forcing of the 6-blade Rushton
impeller, one fused collide-stream step.  Geometry (lattice units, dx = 1):
tank T = H = 480 lu at 512^3 (ratios T/n = 15/16), four wall baffles of
width T/10 and thickness T/48, impeller D = C = T/3, tip speed 0.05 lu,
Re_imp = 5000 (nu = (Omega/2 pi) D^2 / Re).  Inputs are synthetic and
deterministic; the flow starts from rest.

`value` is fluid-cell MLUPS of the whole step with all state resident in
HBM, over a span of steps launched without host synchronisation (the IBM
diagnostics stay on the device).  `roofline` prices the main collide-stream
kernel alone: algorithmic bytes 2 Q S per fluid cell it updates, divided by
its mean launch time from HIP events on the library stream over the same
timed span; `traffic` is the measured HBM bytes per launch from the
committed rocprofv3 counter summary for this workload, if there is one.
`storage_sweep` repeats the timing with the other population storages and
compares the 128^3 tank's power number with each storage against the FP64
oracle series in tests/golden/.  `cpu_baseline` times the CPU oracle (a
restatement of the reference's CpuScalar path) on the same tank at a
reduced size, on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STORAGE_BYTES = {"f32": 4, "f16": 2, "bf16": 2}


def tank_config(n):
    T = 480.0 * n / 512.0
    D = T / 3.0
    omega = 0.05 / (0.5 * D)
    nu = (omega / (2.0 * np.pi)) * D * D / 5000.0
    return {"n": n, "T": T, "H": T, "baffle_width": T / 10.0, "baffle_thickness": T / 48.0, "D": D,
            "clearance": T / 3.0, "omega": omega, "nu": nu}


def profile_name(size, Q, collision, storage):
    return os.path.join(ROOT, "profiles", f"k_hydro_tank{size}_d3q{Q}_{collision}_{storage}.txt")


def profiled_traffic(path):
    """HBM bytes per main-kernel launch from a committed rocprofv3 summary:
    FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the same
    counter reads for a copy kernel of known size in the same run."""
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        return None
    return None


def build_gpu(cfg, storage, collision, Q):
    import lbmflow_amd as L

    n = cfg["n"]
    s = L.Solver(Q, (n, n, n), nu=cfg["nu"], collision=collision, storage=storage)
    s.build_stirred_tank(1.0, cfg["T"], cfg["H"], 4, cfg["baffle_width"], cfg["baffle_thickness"])
    pos, w, vel, c = L.radial_impeller((n, n, n), 1.0, cfg["T"], cfg["H"], cfg["D"], cfg["clearance"], 6, 0.0,
                                       cfg["omega"])
    s.set_rotating_body(pos, w, c, (0.0, 0.0, cfg["omega"]))
    return s, len(w)


def cpu_baseline(n_cpu, steps, collision, Q):
    """The oracle (CpuScalar restatement) on the same tank at n_cpu^3."""
    from oracle import oracle as O

    cfg = tank_config(n_cpu)
    dims = (n_cpu,) * 3
    solid = O.tank_mask(dims, 1.0, cfg["T"], cfg["H"])
    solid, _ = O.baffle_mask(dims, 1.0, cfg["T"], cfg["H"], cfg["baffle_width"], cfg["baffle_thickness"], solid)
    pos, w, vel, c = O.radial_impeller(dims, 1.0, cfg["T"], cfg["H"], cfg["D"], cfg["clearance"], 6, 0.0,
                                       cfg["omega"])
    coll = {"cm": O.CM, "trt": O.TRT, "bgk": O.BGK}[collision]
    orc = O.OracleSolver(Q, dims, prec=4, collision=coll, nu=cfg["nu"], solid=solid)
    fluid = orc.fluid_cells()

    def one():
        orc.clear_force_field()
        orc.apply_ibm(pos, w, c, (0.0, 0.0, cfg["omega"]))
        orc.step()

    one()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    dt = time.perf_counter() - t0
    return fluid * steps / dt / 1e6, dt


def timed_span(s, steps):
    """`steps` runner steps with no host synchronisation inside; HIP events
    on the library stream bracket every main-kernel launch."""
    s.synchronize()
    s.kernel_timing(True)
    t0 = time.perf_counter()
    tq, diag = s.run_with_ibm(steps, graph=False)
    s.synchronize()
    dt = time.perf_counter() - t0
    k_ms, k_n = s.kernel_timing_read()
    s.kernel_timing(False)
    return dt, k_ms / max(k_n, 1), tq, diag


def np_error_vs_f64(storages, collision, Q):
    """Np of the 128^3 tank (last-half mean torque, runner.rs:1514-1530)
    against the FP64 oracle series stored in tests/golden/."""
    import lbmflow_amd as L

    path = os.path.join(ROOT, "tests", "golden", "tank128_cm_f64.npz")
    if not os.path.exists(path) or collision != "cm" or Q != 27:
        return None
    z = np.load(path)
    meta = json.loads(bytes(z["meta"]).decode())
    steps = int(meta["steps"])
    cfg = tank_config(meta["n"])
    rec = np.arange(1, steps + 1)

    def np_of(torque_z):
        tq = float(np.mean(torque_z[rec >= steps // 2]))
        return L.power_qois(tq, cfg["omega"], 1.0, cfg["D"], 1.0)["np"]

    ref = np_of(np.asarray(z["torque"]))
    out = {"oracle": "f64 CPU restatement, 128^3, %d steps" % steps, "np_f64": ref}
    for st in storages:
        s, _ = build_gpu(cfg, st, collision, Q)
        tq, _ = s.run_with_ibm(steps, graph=False)
        s.close()
        v = np_of(-tq[:, 2])
        out[st] = {"np": v, "rel_err": abs(v - ref) / abs(ref)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--collision", default="cm", choices=["cm", "trt", "bgk"])
    ap.add_argument("--lattice", type=int, default=27, choices=[19, 27])
    ap.add_argument("--cpu-size", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sweep", action="store_true", help="skip the 16-bit storage and Np-accuracy sweep")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is not built",
              file=sys.stderr)
        sys.exit(2)

    cfg = tank_config(args.size)
    t_setup = time.perf_counter()
    s, n_markers = build_gpu(cfg, args.storage, args.collision, args.lattice)
    fluid = s.fluid_cells()
    n_total = args.size ** 3
    n_list = s.force_list_size()
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] setup {setup_s:.1f}s  n={args.size}  fluid={int(fluid)}  markers={n_markers}  "
          f"force-list={n_list}  device={s.device_bytes() / 2**30:.2f} GiB", file=sys.stderr, flush=True)

    s.run_with_ibm(args.warmup, graph=False)
    dt, k_avg_ms, tq, diag = timed_span(s, args.steps)
    mass = s.total_mass_f64()

    Q = args.lattice
    S = STORAGE_BYTES[args.storage]
    B = 2 * Q * S
    ms_per_step = 1e3 * dt / args.steps
    mlups_fluid = fluid * args.steps / dt / 1e6
    mlups_total = n_total * args.steps / dt / 1e6
    cells_main = fluid - n_list
    achieved = cells_main * B / (k_avg_ms * 1e-3) / 1e9
    print(f"[bench] {args.storage}: {ms_per_step:.3f} ms/step, k_hydro {k_avg_ms:.3f} ms, "
          f"{achieved:.0f} GB/s algorithmic", file=sys.stderr, flush=True)

    copy_gbps = s.ctx.stream_copy_gbps(4 << 30, 20)
    prof = profile_name(args.size, args.lattice, args.collision, args.storage)
    traffic = profiled_traffic(prof)
    s.close()

    sweep = None
    if not args.no_sweep:
        sweep = {}
        for st in [v for v in ("f32", "f16", "bf16") if v != args.storage]:
            s2, _ = build_gpu(cfg, st, args.collision, Q)
            s2.run_with_ibm(200, graph=False)
            dt2, k2, _, _ = timed_span(s2, 100)
            s2.close()
            sweep[st] = {"mlups_fluid": round(fluid * 100 / dt2 / 1e6, 2), "kernel_avg_ms": round(k2, 4),
                         "speedup_vs_" + args.storage: round(dt / args.steps / (dt2 / 100), 3)}
            print(f"[bench] {st}: {1e3 * dt2 / 100:.3f} ms/step, k_hydro {k2:.3f} ms", file=sys.stderr, flush=True)
        nperr = np_error_vs_f64(("f32", "f16", "bf16"), args.collision, Q)
        if nperr is not None:
            sweep["np_vs_f64_oracle"] = nperr

    cpu = None
    if not args.no_cpu_baseline:
        threads = os.cpu_count() or 1
        omp = os.environ.get("OMP_NUM_THREADS")
        if omp:
            threads = int(omp)
        v, secs = cpu_baseline(args.cpu_size, args.cpu_steps, args.collision, Q)
        cpu = {"value": round(v, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
               "sample": f"oracle D3Q{Q} {args.collision} fp32 stirred tank {args.cpu_size}^3 (same ratios), "
                         f"{args.cpu_steps} steps incl. IBM, {secs:.1f}s"}

    out = {
        "metric": "MLUPS & % HBM roofline, D3Q27 tank 512^3",
        "value": round(mlups_fluid, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"f32 math, {args.storage} storage",
        "data": "synthetic (deterministic tank geometry and Rushton IBM markers, flow from rest)",
        "config": {"workload": f"stirred tank D3Q{Q} {args.collision} {args.size}^3 + rotating IBM",
                   "lattice": f"D3Q{Q}", "collision": args.collision, "storage": args.storage,
                   "grid": [args.size] * 3, "tank_T_lu": cfg["T"], "impeller_D_lu": cfg["D"],
                   "tip_speed_lu": 0.05, "re_imp": 5000, "nu_lu": cfg["nu"], "markers": n_markers,
                   "fluid_cells": int(fluid), "parallelism": "single GPU"},
        "mlups_total": round(mlups_total, 2),
        "mlups_fluid": round(mlups_fluid, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_source": os.path.relpath(prof, ROOT) if traffic is not None else None,
                     "algorithmic_bytes_per_launch": int(cells_main * B),
                     "kernel": "k_hydro (fused AA collide-stream)", "kernel_avg_ms": round(k_avg_ms, 4),
                     "bytes_per_cell": B, "cells_per_launch": int(cells_main),
                     "stream_copy_gbs": round(copy_gbps, 1), "frac_of_copy": round(achieved / copy_gbps, 4)},
        "storage_sweep": sweep,
        "cpu_baseline": cpu,
        "checks": {"mass_finite": bool(np.isfinite(mass)), "mass_total": mass,
                   "torque_z": float(diag["torque"][2]), "ibm_iterations": diag["iterations"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
