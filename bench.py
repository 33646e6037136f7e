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
restatement of the reference's CpuScalar path) on the same tank, a short
window from the GPU run's developed state, on this host's cores.
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


def cpu_info():
    """The host the CPU leg runs on: model, logical CPUs, OMP_NUM_THREADS."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(cfg, f_dev, gpu_steps, steps, collision, Q):
    """The oracle (the CpuScalar restatement in fp32, OpenMP over rows) on
    the benchmark's own tank, started from the populations the GPU run
    reached after `gpu_steps` steps: a flow from rest exercises no storage
    rounding and converges the IBM in fewer sweeps.  One untimed step, then
    `steps` timed runner steps (clear force, rotating IBM, collide-stream)."""
    from oracle import oracle as O

    n = cfg["n"]
    dims = (n,) * 3
    solid = O.tank_mask(dims, 1.0, cfg["T"], cfg["H"])
    solid, _ = O.baffle_mask(dims, 1.0, cfg["T"], cfg["H"], cfg["baffle_width"], cfg["baffle_thickness"], solid)
    pos, w, vel, c = O.radial_impeller(dims, 1.0, cfg["T"], cfg["H"], cfg["D"], cfg["clearance"], 6, 0.0,
                                       cfg["omega"])
    coll = {"cm": O.CM, "trt": O.TRT, "bgk": O.BGK}[collision]
    orc = O.OracleSolver(Q, dims, prec=4, collision=coll, nu=cfg["nu"], solid=solid)
    fluid = orc.fluid_cells()
    orc.set_f(f_dev)
    orc.update_moments()

    def one():
        orc.clear_force_field()
        orc.apply_ibm(pos, w, c, (0.0, 0.0, cfg["omega"]))
        orc.step()

    one()
    t0 = time.perf_counter()
    for k in range(steps):
        one()
        print(f"[bench] cpu oracle {n}^3: step {k + 1}/{steps}", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    info = dict(cpu_info(), threads_used=O.max_threads())
    return fluid * steps / dt / 1e6, dt, info


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


def qoi_error_vs_f64(storages, collision, Q):
    """The 128^3 tank's QOIs with each population storage against the FP64
    oracle series in tests/golden/ (Re_imp 100, D3Q7 tracer with Sc = 1, the
    fixture's full step count): Np from the last-half mean shaft torque
    (runner.rs:1514-1530, qoi.rs:477-505), here the device's running FP64
    torque integral over that window, and the 95 % mixing time from the
    tracer CV sampled every cv_every steps (qoi.rs:534-556)."""
    import lbmflow_amd as L

    path = os.path.join(ROOT, "tests", "golden", "tank128_cm_f64.npz")
    if not os.path.exists(path) or collision != "cm" or Q != 27:
        return None
    z = np.load(path)
    meta = json.loads(bytes(z["meta"]).decode())
    steps, every, n = int(meta["steps"]), int(meta["cv_every"]), int(meta["n"])
    cfg = dict(tank_config(n), nu=meta["nu"])
    rec = np.arange(1, steps + 1)
    t_cv = every * np.arange(1, steps // every + 1, dtype=float)
    tq_ref = np.asarray(z["torque"])
    np_ref = L.power_qois(float(np.mean(tq_ref[rec >= steps // 2])), cfg["omega"], 1.0, cfg["D"], 1.0)["np"]
    t95_ref = L.mixing_time_from_cv(t_cv, np.asarray(z["cv"]))["t95"]
    out = {"oracle": f"f64 CPU restatement, {n}^3, Re_imp {meta['re_imp']:g}, {steps} steps, D3Q7 tracer",
           "np_f64": np_ref, "t95_f64": t95_ref}
    w0 = steps // 2 - 1  # the torque window is steps steps//2 .. steps
    for st in storages:
        s, _ = build_gpu(cfg, st, collision, Q)
        solid = s.solid().reshape(n, n, n)
        zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        sig, cz = float(meta["sigma"]), 0.8 * cfg["H"]
        blob = np.exp(-((xx + 0.5 - n / 2) ** 2 + (yy + 0.5 - n / 2) ** 2 + (zz + 0.5 - cz) ** 2) / (2 * sig ** 2))
        blob = np.where(solid == 0, blob, 0.0).astype(np.float32).reshape(-1)
        k = s.enable_scalar_distribution(cfg["nu"], "d3q7")
        s.set_scalar_concentration(k, blob)
        cv = []
        done = 0
        while done < steps:
            step = min(every - done % every, w0 - done if done < w0 else every)
            s.run_with_ibm(step, graph=True)
            done += step
            if done == w0:
                s.torque_integral(reset=True)
            if done % every == 0:
                cv.append(s.scalar_cv(k))
        tsum, cnt = s.torque_integral()
        s.close()
        v = L.power_qois(float(-tsum[2] / cnt), cfg["omega"], 1.0, cfg["D"], 1.0)["np"]
        t95 = L.mixing_time_from_cv(t_cv, np.array(cv))["t95"]
        out[st] = {"np": v, "np_rel_err": abs(v - np_ref) / abs(np_ref), "t95": t95,
                   "t95_rel_err": None if t95 is None or t95_ref is None else abs(t95 - t95_ref) / t95_ref}
        print(f"[bench] {st} vs f64: Np rel {out[st]['np_rel_err']:.2e}, t95 {t95} vs {t95_ref}",
              file=sys.stderr, flush=True)
    return out


def splitmix64(seed):
    """Deterministic uniform draws in [0, 1) (SplitMix64)."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        yield (z >> 11) / float(1 << 53)


def small_tank(n, tip, nu, Q=19, collision="trt"):
    """A tank at the benchmark's ratios on an n^3 grid, tip speed `tip`."""
    import lbmflow_amd as L

    T = 15.0 * n / 16.0
    D = T / 3.0
    omega = tip / (0.5 * D)
    s = L.Solver(Q, (n, n, n), nu=nu, collision=collision)
    s.build_stirred_tank(1.0, T, T, 4, T / 10.0, T / 48.0)
    pos, w, vel, c = L.radial_impeller((n, n, n), 1.0, T, T, D, T / 3.0, 6, 0.0, omega)
    s.set_rotating_body(pos, w, c, (0.0, 0.0, omega))
    return s, {"n": n, "T": T, "D": D, "omega": omega, "tip": tip, "nu": nu, "markers": len(w)}


def workload_precision512(args):
    """The precision configuration at its full size: the 512^3 tank (D3Q27
    central moments, Re_imp 5000, per-step WALE: without a turbulence model
    this tank diverges after about 3,150 steps) with --storage population
    storage and a D3Q7 tracer pulse (Gaussian sigma 8 lu at 0.8 H on the
    axis, Sc = 1), run from rest until the 95 % mixing time (CV sampled every
    100 steps, qoi.rs:534-556) or --steps; Np from the last-half mean shaft
    torque (runner.rs:1514-1530).  The errors against the FP64 oracle are
    measured at 128^3 (qoi_vs_f64_oracle of the default workload); this run
    gives the storages' Np and t95 at 512^3 side by side."""
    import lbmflow_amd as L

    n = 512
    cfg = tank_config(n)
    s, n_markers = build_gpu(cfg, args.storage, "cm", 27)
    s.set_wale(True)
    fluid = s.fluid_cells()
    solid = s.solid().reshape(n, n, n)
    ax = np.arange(n, dtype=np.float64) + 0.5 - n / 2
    zc = np.arange(n, dtype=np.float64) + 0.5 - 0.8 * cfg["H"]
    sig = 8.0
    blob = np.exp(-(ax[None, None, :] ** 2 + ax[None, :, None] ** 2 + zc[:, None, None] ** 2) / (2 * sig * sig))
    blob = np.where(solid == 0, blob, 0.0).astype(np.float32).reshape(-1)
    del solid
    k = s.enable_scalar_distribution(cfg["nu"], "d3q7")
    s.set_scalar_concentration(k, blob)
    del blob
    every, torque, t_cv, cv = 100, [], [], []
    t95, done = None, 0
    s.synchronize()
    t0 = time.perf_counter()
    while done < args.steps:
        tq, diag = s.run_with_ibm(every, graph=True)
        done += every
        torque.append(-tq[:, 2])
        c = s.scalar_cv(k)
        t_cv.append(float(done))
        cv.append(c)
        if not np.isfinite(c) or not np.isfinite(tq).all():
            break
        if done % 1000 == 0:
            print(f"[bench] precision512 {args.storage}: {done} steps, CV {c:.4f}", file=sys.stderr, flush=True)
        mix = L.mixing_time_from_cv(np.array(t_cv), np.array(cv))
        if mix is not None and mix["t95"] is not None:
            t95 = mix["t95"]
            break
    s.synchronize()
    dt = time.perf_counter() - t0
    torque = np.concatenate(torque)
    finite = bool(np.isfinite(torque).all() and np.isfinite(cv).all())
    tmean = float(np.mean(torque[len(torque) // 2:]))
    q = L.power_qois(tmean, cfg["omega"], 1.0, cfg["D"], 1.0)
    s.close()
    rev = done * cfg["omega"] / (2 * np.pi)
    return {
        "metric": "Np and 95 % mixing time, D3Q27 CM tank 512^3 with WALE and a D3Q7 tracer",
        "value": round(fluid * done / dt / 1e6, 2), "unit": "MLUPS", "n_gpus": 1, "steps": done, "warmup": 0,
        "ms_per_step": round(1e3 * dt / done, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.storage == "f32" else f"f32 math, {args.storage} storage",
        "data": "synthetic (tank geometry, Rushton lattice markers, Gaussian tracer sigma 8 lu; flow from rest)",
        "config": {"workload": "stirred tank D3Q27 cm 512^3 + per-step WALE + D3Q7 tracer (Sc 1), until t95",
                   "storage": args.storage, "grid": [n] * 3, "re_imp": 5000, "nu_lu": cfg["nu"],
                   "markers": n_markers, "fluid_cells": int(fluid), "cv_every": every,
                   "parallelism": "single GPU"},
        "np": q["np"], "t95_steps": t95, "t95_revolutions": None if t95 is None else t95 * cfg["omega"] / (2 * np.pi),
        "revolutions_run": rev, "cv0": cv[0] if cv else None, "cv_last": cv[-1] if cv else None,
        "checks": {"finite": finite},
    }


def workload_calib(args):
    """Counter calibration, not a throughput benchmark: one GiB copied with
    each lane access form of the collide-stream kernels (16-byte; 4-byte
    aligned and shifted one element, as the odd step's x-neighbour rows;
    2-byte aligned and shifted), `steps` launches each, then the 4 GiB
    16-byte stream copy.  Under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE the
    per-dispatch counts divided by the bytes below give each form's factor."""
    import lbmflow_amd as L

    ctx = L.default_context()
    nbytes = 1 << 30
    forms = [(16, 0), (4, 0), (4, 1), (2, 0), (2, 1)]
    out = {}
    for w, o in forms:
        out[f"w{w}_off{o}"] = round(ctx.access_copy_gbps(w, o, nbytes, args.steps), 1)
        print(f"[bench] access copy width {w} offset {o}: {out[f'w{w}_off{o}']} GB/s", file=sys.stderr, flush=True)
    out["stream_copy_4GiB"] = round(ctx.stream_copy_gbps(4 << 30, args.steps), 1)
    return {"metric": "copy GB/s per lane access form (counter calibration)", "value": out["w16_off0"],
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic (zeros)",
            "config": {"workload": "access-form copies, 1 GiB read + 1 GiB written per launch",
                       "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                       "kernels": {"k_access_copy<uint4, 0>": "w16_off0", "k_access_copy<unsigned int, 0>": "w4_off0",
                                   "k_access_copy<unsigned int, 1>": "w4_off1",
                                   "k_access_copy<unsigned short, 0>": "w2_off0",
                                   "k_access_copy<unsigned short, 1>": "w2_off1"}},
            "gbps": out}


def workload_cavity(args):
    """The plain-method configuration: 3D lid-driven cavity, D3Q19 TRT
    (magic 3/16), 64^3, fp32, Re 100, lid 0.05 lu, quiescent start, 1000
    steps; MLUPS and the centreline velocity profiles of the reference's
    cavity test (t15_5_cavity3d.rs:183-243)."""
    import lbmflow_amd as L
    from lbmflow_amd import cavity

    n, u_lid, re = args.size if args.size != 512 else 64, 0.05, 100.0
    nu = u_lid * (n - 2) / re
    s = L.Solver(19, (n, n, n), nu=nu, collision="trt")
    u = np.zeros((6, 3))
    u[5] = [u_lid, 0.0, 0.0]
    s.build_wall_rims([1] * 6, u)
    fluid = s.fluid_cells()
    s.synchronize()
    t0 = time.perf_counter()
    s.run(args.steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    rho, ux, uy, uz = s.moments()
    zf, uu, xf, ww = cavity.centrelines(ux, uz, n, u_lid)
    mass = s.mass_deviation()
    s.close()
    return {
        "metric": "MLUPS, 3D lid-driven cavity D3Q19 TRT 64^3 Re 100 (1000 steps)",
        "value": round(n ** 3 * args.steps / dt / 1e6, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": 0, "ms_per_step": round(1e3 * dt / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (quiescent start)",
        "config": {"workload": f"lid-driven cavity D3Q19 TRT {n}^3, Re 100, lid 0.05", "grid": [n] * 3,
                   "fluid_cells": int(fluid), "parallelism": "single GPU"},
        "mlups_fluid": round(fluid * args.steps / dt / 1e6, 2),
        "centreline_u_vs_z": [[float(a), float(b)] for a, b in zip(zf, uu)],
        "centreline_w_vs_x": [[float(a), float(b)] for a, b in zip(xf, ww)],
        "checks": {"mass_deviation": float(mass)},
    }


def workload_surface(args):
    """The single-phase stirred tank at 256^3 (D3Q27 central moment, fp32,
    T = H = 240 lu, 4 baffles T/10, Rushton D = C = T/3, tip speed 0.05,
    Re_imp 5000) with the seeded surface sampling of the blades (one marker
    per 0.1 lu^2: about 12k markers), 20k steps; Np from the last-half mean
    shaft torque (runner.rs:1514-1546, qoi.rs:477-505).  The WALE closure
    runs before every step (les.rs): without a turbulence model this tank
    (nu = 2.5e-4 lu) diverges after about 2,400 steps."""
    import lbmflow_amd as L

    n = 256
    cfg = tank_config(n)
    s = L.Solver(27, (n, n, n), nu=cfg["nu"], collision="cm", storage=args.storage)
    s.build_stirred_tank(1.0, cfg["T"], cfg["H"], 4, cfg["baffle_width"], cfg["baffle_thickness"])
    pos, w, vel, c = L.impeller_surface_markers((n, n, n), 1.0, cfg["T"], cfg["H"], cfg["D"], cfg["clearance"], 6,
                                                0.0, cfg["omega"], 0.04, 0.1, args.seed)
    s.set_rotating_body(pos, w, c, (0.0, 0.0, cfg["omega"]))
    s.set_wale(True)
    fluid = s.fluid_cells()
    steps = args.steps
    torque = []
    chunk = 1000
    # the QOI window is steps steps//2 .. steps (1-based): the device's FP64
    # torque integral is cleared after step steps//2 - 1
    w0 = max(steps // 2 - 1, 0)
    s.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(chunk, (w0 if done < w0 else steps) - done)
        tq, diag = s.run_with_ibm(k, graph=True)
        torque.append(-tq[:, 2])
        done += k
        if done == w0:
            s.torque_integral(reset=True)
        print(f"[bench] surface tank: {done}/{steps} steps", file=sys.stderr, flush=True)
    s.synchronize()
    dt = time.perf_counter() - t0
    tsum, tcount = s.torque_integral()
    tmean = float(-tsum[2] / tcount)
    torque = np.concatenate(torque)
    rec = np.arange(1, steps + 1)
    tmean_host = float(np.mean(torque[rec >= steps // 2]))
    q = L.power_qois(tmean, cfg["omega"], 1.0, cfg["D"], 1.0)
    mass = s.total_mass_f64()
    s.close()
    return {
        "metric": "MLUPS and Np, D3Q27 CM stirred tank 256^3 with surface-sampled Rushton markers",
        "value": round(fluid * steps / dt / 1e6, 2), "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(1e3 * dt / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.storage == "f32" else f"f32 math, {args.storage} storage",
        "data": f"synthetic (tank geometry; surface markers SplitMix64 seed {args.seed}; flow from rest)",
        "config": {"workload": "stirred tank D3Q27 cm 256^3 + per-step WALE, surface-sampled Rushton IBM",
                   "grid": [n] * 3,
                   "markers": int(len(w)), "marker_area_lu2": 0.1, "tip_speed_lu": 0.05, "re_imp": 5000,
                   "nu_lu": cfg["nu"], "fluid_cells": int(fluid), "parallelism": "single GPU"},
        "np": q["np"], "torque_mean_lu": tmean, "revolutions": steps * cfg["omega"] / (2 * np.pi),
        "checks": {"mass_finite": bool(np.isfinite(mass)), "ibm_iterations_last": diag["iterations"],
                   "window_updates": tcount, "torque_mean_host_series": tmean_host,
                   "torque_mean_rel_diff": abs(tmean - tmean_host) / max(abs(tmean_host), 1e-300)},
    }


def workload_batch(args):
    """64 tanks in one persistent launch (edges U[48,160], tip U[0.02,0.06],
    nu U[0.002,0.02], D3Q19 TRT fp32), graph-captured step, autotuned grid."""
    import lbmflow_amd as L

    rng = splitmix64(args.seed)
    specs = [(48 + int(next(rng) * 113), 0.02 + 0.04 * next(rng), 0.002 + 0.018 * next(rng)) for _ in range(64)]
    t0 = time.perf_counter()
    solvers, meta = [], []
    for n, tip, nu in specs:
        sv, m = small_tank(n, tip, nu)
        solvers.append(sv)
        meta.append(m)
    batch = L.Batch(solvers)
    grid, trial_ms = batch.autotune(10)
    setup_s = time.perf_counter() - t0
    fluid = sum(sv.fluid_cells() for sv in solvers)
    total = sum(m["n"] ** 3 for m in meta)
    print(f"[bench] batch setup {setup_s:.1f}s  tanks=64  cells={total}  fluid={int(fluid)}  grid={grid}",
          file=sys.stderr, flush=True)
    batch.run(args.warmup, graph=True)
    solvers[0].synchronize()
    # two spans: the second is the QOI window, over which the device's FP64
    # torque integrals run (reset between the spans)
    h = args.steps // 2
    t0 = time.perf_counter()
    tq1 = batch.run(h, graph=True)
    batch.torque_integrals(reset=True)
    tq2 = batch.run(args.steps - h, graph=True)
    solvers[0].synchronize()
    dt = time.perf_counter() - t0
    tsum, tcnt = batch.torque_integrals()
    tq = np.concatenate([tq1, tq2], axis=1)
    nps, dev = [], 0.0
    for k, m in enumerate(meta):
        tz = -tq[k, :, 2]
        t_host = float(np.mean(tz[h:]))
        tmean = float(-tsum[k, 2] / tcnt[k])
        dev = max(dev, abs(tmean - t_host) / max(abs(t_host), 1e-300))
        nps.append(L.power_qois(tmean, m["omega"], 1.0, m["D"], 1.0)["np"])
    batch.close()
    B = 2 * 19 * 4
    achieved = fluid * B / (trial_ms * 1e-3) / 1e9
    return {
        "metric": "aggregate MLUPS, 64-tank batched sweep (D3Q19 TRT fp32, one persistent launch)",
        "value": round(fluid * args.steps / dt / 1e6, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic (SplitMix64 seed {args.seed}: edges U[48,160], tip U[0.02,0.06], nu U[0.002,0.02])",
        "config": {"workload": "64 stirred tanks, D3Q19 TRT, rotating IBM, hipGraph step", "tanks": 64,
                   "cells": int(total), "fluid_cells": int(fluid), "persistent_grid": grid,
                   "autotune_trial_ms": round(trial_ms, 4), "parallelism": "single GPU"},
        "mlups_total": round(total * args.steps / dt / 1e6, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_hydro_batch (persistent collide-stream over all tanks)",
                     "kernel_avg_ms": round(trial_ms, 4), "bytes_per_cell": B,
                     "algorithmic_bytes_per_launch": int(fluid * B),
                     "timing": "mean launch time of the autotuner's best grid (HIP events, 10 launches)"},
        "checks": {"np_window_steps": args.steps - h, "device_vs_host_torque_mean_rel": dev},
        "per_instance": [{"n": m["n"], "tip": round(m["tip"], 5), "nu": round(m["nu"], 6), "np": round(v, 5)}
                         for m, v in zip(meta, nps)],
    }


def workload_adjoint(args):
    """Adjoint of the time-averaged torque: 128^3 D3Q19 TRT tank, nu = 0.01
    +-10 % smooth seeded field, 4000 steps, window 1000, binomial
    checkpointing with 128 snapshots (24 GB of HBM; the recomputation drops
    from 15.5 k forward steps with 64 to 11.9 k, and more snapshots cost more
    in allocation than they save); checked by central differences."""
    n, steps, window, snaps = 128, args.steps, args.steps // 4, args.snapshots
    rng = splitmix64(args.seed)
    ph = [2 * np.pi * next(rng) for _ in range(3)]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
    field = np.sin(2 * np.pi * x / n + ph[0]) * np.cos(2 * np.pi * y / n + ph[1]) * np.sin(2 * np.pi * z / n + ph[2])
    nu = 0.01 * (1.0 + 0.1 * field)

    def make(omega_scale=1.0, nuf=nu):
        s, m = small_tank(n, 0.05 * omega_scale, 0.01)
        s.set_omega_field((1.0 / (3.0 * nuf + 0.5)).astype(np.float32))
        return s, m

    s, m = make()
    t0 = time.perf_counter()
    res = s.adjoint_torque(steps, window, snaps)
    s.synchronize()
    dt = time.perf_counter() - t0
    solid = s.solid()
    s.close()

    def J(scale=1.0, nuf=nu):
        sv, _ = make(scale, nuf)
        v = sv.adjoint_torque(steps, window, 0, gradients=False)["J"]
        sv.close()
        return v

    eps = 1e-2
    fd_w = (J(1 + eps) - J(1 - eps)) / (2 * eps * m["omega"])
    d = 0.01 * np.cos(2 * np.pi * (x + 2 * y) / n) * np.sin(2 * np.pi * z / n) * (solid == 0)
    fd_nu = (J(1.0, nu + eps * d) - J(1.0, nu - eps * d)) / (2 * eps)
    ad_nu = float(np.dot(res["dJ_dnu"], d))
    return {
        "metric": "torque adjoint: time for forward + reverse, 128^3 D3Q19 TRT tank with rotating IBM",
        "value": round(dt, 3), "unit": "s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(1e3 * dt / steps, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 forward, f64 adjoint",
        "data": f"synthetic (tank ratios of the benchmark, nu field seed {args.seed})",
        "config": {"workload": "adjoint of mean shaft torque, window %d, %d binomial snapshots" % (window, snaps),
                   "grid": [n] * 3, "forward_steps_total": res["forward_steps"], "parallelism": "single GPU"},
        "J": res["J"],
        "dJ_domega": {"adjoint": res["dJ_domega"], "central_fd_fp32": fd_w,
                      "rel_diff": abs(res["dJ_domega"] - fd_w) / abs(fd_w)},
        "dJ_dnu_directional": {"adjoint": ad_nu, "central_fd_fp32": fd_nu,
                               "rel_diff": abs(ad_nu - fd_nu) / abs(fd_nu)},
    }


def workload_channel(args):
    """Open channel: D3Q19 TRT fp32, 768 x 256 x 256, Zou-He velocity inlet
    (x-), Zou-He pressure outlet (x+), half-way walls at y = 0 and ny - 1,
    periodic z.  Reports the step rate, the collide-stream kernel's HBM
    rate and what the open-face passes (capture, restore, two face rules)
    add per step."""
    import lbmflow_amd as L

    dims = (768, 256, 256)
    nx, ny, nz = dims
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    s = L.Solver(19, dims, nu=0.02, collision="trt", periodic=(False, False, True), solid=solid.reshape(-1))
    s.set_face_bc("x-", "velocity", (0.05, 0.0, 0.0))
    s.set_face_bc("x+", "pressure", (1.0,))
    fluid = s.fluid_cells()
    s.run(args.warmup)
    s.synchronize()
    s.kernel_timing(True)
    t0 = time.perf_counter()
    s.run(args.steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    k_ms, k_n = s.kernel_timing_read()
    s.kernel_timing(False)
    k_avg = k_ms / max(k_n, 1)
    mass = s.total_mass_f64()
    copy_gbps = s.ctx.stream_copy_gbps(4 << 30, 20)
    s.close()
    ms = 1e3 * dt / args.steps
    B = 2 * 19 * 4
    achieved = fluid * B / (k_avg * 1e-3) / 1e9
    print(f"[bench] channel: {ms:.3f} ms/step, k_hydro {k_avg:.3f} ms, {achieved:.0f} GB/s", file=sys.stderr,
          flush=True)
    return {
        "metric": "MLUPS, D3Q19 TRT open channel 768x256x256 (Zou-He velocity inlet, pressure outlet)",
        "value": round(fluid * args.steps / dt / 1e6, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (channel from rest)",
        "config": {"workload": "open channel D3Q19 TRT 768x256x256, velocity inlet / pressure outlet, z periodic",
                   "grid": list(dims), "fluid_cells": int(fluid), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_hydro (fused AA collide-stream)", "kernel_avg_ms": round(k_avg, 4),
                     "bytes_per_cell": B, "stream_copy_gbs": round(copy_gbps, 1)},
        "open_face_ms_per_step": round(ms - k_avg, 4),
        "checks": {"mass_finite": bool(np.isfinite(mass))},
    }


def workload_les(args):
    """The 512^3 tank step with the WALE closure recomputed on the device
    before every collision (moments, WALE omega field, collide-stream with
    that field); reports the step rate and the closure's share of it."""
    cfg = tank_config(args.size)
    Q = args.lattice
    s, n_markers = build_gpu(cfg, args.storage, args.collision, Q)
    fluid = s.fluid_cells()
    s.run_with_ibm(args.warmup, graph=False)
    dt0, k0, _, _ = timed_span(s, args.steps)
    s.set_wale(True)
    s.run_with_ibm(10, graph=False)
    dt, k_ms, tq, diag = timed_span(s, args.steps)
    d, nut = s.wale_update(want_nu_t=True)
    copy_gbps = s.ctx.stream_copy_gbps(4 << 30, 20)
    s.close()
    ms, ms0 = 1e3 * dt / args.steps, 1e3 * dt0 / args.steps
    S = STORAGE_BYTES[args.storage]
    # the closure's algorithmic bytes: per fluid cell the local pass reads Q
    # populations and the mask and writes the velocity and 3 strain
    # components, the gradient pass reads the mask, the cell's velocity and
    # strain and writes omega (neighbour velocities come from L2); per solid
    # cell the local pass reads the mask and writes 6 zeros, the gradient pass
    # reads the mask and writes omega
    n_total = args.size ** 3
    b_cell = (Q * S + 4 + 24) + (4 + 24 + 4)
    b_solid = (4 + 24) + (4 + 4)
    closure_bytes = fluid * b_cell + (n_total - fluid) * b_solid
    wale_ms = ms - ms0
    print(f"[bench] les: {ms:.3f} ms/step (plain {ms0:.3f}), closure {wale_ms:.3f} ms, "
          f"max nu_t {d['max_nu_t_before_clipping']:.3e}", file=sys.stderr, flush=True)
    return {
        "metric": "MLUPS, D3Q27 tank 512^3 with per-step WALE LES",
        "value": round(fluid * args.steps / dt / 1e6, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.storage == "f32" else f"f32 math, {args.storage} storage",
        "data": "synthetic (deterministic tank geometry and Rushton IBM markers, flow from rest)",
        "config": {"workload": f"stirred tank D3Q{Q} {args.collision} {args.size}^3 + rotating IBM + WALE per step",
                   "grid": [args.size] * 3, "fluid_cells": int(fluid), "markers": n_markers,
                   "parallelism": "single GPU"},
        "plain_ms_per_step": round(ms0, 4),
        "closure_ms_per_step": round(wale_ms, 4),
        "closure_algorithmic_gbs": round(closure_bytes / (wale_ms * 1e-3) / 1e9, 1) if wale_ms > 0 else None,
        "closure_bytes_per_fluid_cell": b_cell,
        "closure_bytes_per_solid_cell": b_solid,
        "stream_copy_gbs": round(copy_gbps, 1),
        "k_hydro_avg_ms": round(k_ms, 4),
        "max_nu_t": d["max_nu_t_before_clipping"],
        "mean_nu_t_over_nu": float(np.mean(nut[nut > 0]) / cfg["nu"]) if np.any(nut > 0) else 0.0,
        "checks": {"torque_z": float(diag["torque"][2])},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="default 200 (tank), 5000 (batch), 4000 (adjoint)")
    ap.add_argument("--warmup", type=int, default=None, help="default 500 (tank), 100 (batch), 0 (adjoint)")
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--collision", default="cm", choices=["cm", "trt", "bgk"])
    ap.add_argument("--lattice", type=int, default=27, choices=[19, 27])
    ap.add_argument("--cpu-steps", type=int, default=10, help="timed oracle steps of the tank on the host")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sweep", action="store_true", help="skip the 16-bit storage and Np-accuracy sweep")
    ap.add_argument("--workload", default="tank", choices=["tank", "batch", "adjoint", "les", "channel", "surface",
                                                              "cavity", "calib", "precision512"],
                    help="tank: the 512^3 benchmark (default); batch: 64-tank sweep; adjoint: torque gradient; "
                         "les: the tank with per-step WALE; channel: open-face channel; surface: the 256^3 tank "
                         "with surface-sampled markers, 20k steps, Np; cavity: the 64^3 lid-driven cavity; "
                         "calib: copies per lane access form for counter calibration; precision512: the 512^3 "
                         "tank with WALE and a tracer until t95, per --storage")
    ap.add_argument("--seed", type=int, default=20260512)
    ap.add_argument("--snapshots", type=int, default=128, help="adjoint: binomial checkpoint snapshots")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is not built",
              file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = {"tank": 200, "batch": 5000, "adjoint": 4000, "les": 100, "channel": 200,
                      "surface": 20000, "cavity": 1000, "calib": 5, "precision512": 90000}[args.workload]
    if args.warmup is None:
        args.warmup = {"tank": 500, "batch": 100, "adjoint": 0, "les": 200, "channel": 100, "surface": 0,
                       "cavity": 0, "calib": 1, "precision512": 0}[args.workload]
    if args.workload != "tank":
        fn = {"batch": workload_batch, "adjoint": workload_adjoint, "les": workload_les,
              "channel": workload_channel, "surface": workload_surface,
              "cavity": workload_cavity, "calib": workload_calib,
              "precision512": workload_precision512}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return

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
    # tile visit order per parity, timed on a scratch copy of the state
    tune = s.autotune_tiles(20)
    print(f"[bench] autotuned tile order: {tune}", file=sys.stderr, flush=True)
    dt, k_avg_ms, tq, diag = timed_span(s, args.steps)
    # the same span replayed from the captured two-step HIP graph
    s.run_with_ibm(2, graph=True)  # capture (not timed)
    s.synchronize()
    t0 = time.perf_counter()
    s.run_with_ibm(args.steps, graph=True)
    s.synchronize()
    dt_graph = time.perf_counter() - t0
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
          f"{achieved:.0f} GB/s algorithmic; graph-replayed span {1e3 * dt_graph / args.steps:.3f} ms/step",
          file=sys.stderr, flush=True)

    copy_gbps = s.ctx.stream_copy_gbps(4 << 30, 20)
    prof = profile_name(args.size, args.lattice, args.collision, args.storage)
    traffic = profiled_traffic(prof)
    steps_done = s.time
    f_dev = None if args.no_cpu_baseline else s.populations()
    s.close()

    sweep = None
    if not args.no_sweep:
        sweep = {}
        for st in [v for v in ("f32", "f16", "bf16") if v != args.storage]:
            s2, _ = build_gpu(cfg, st, args.collision, Q)
            s2.run_with_ibm(args.warmup, graph=False)
            tune2 = s2.autotune_tiles(20)
            dt2, k2, _, _ = timed_span(s2, args.steps)
            s2.close()
            sweep[st] = {"mlups_fluid": round(fluid * args.steps / dt2 / 1e6, 2), "kernel_avg_ms": round(k2, 4),
                         "speedup_vs_" + args.storage: round(dt / dt2, 3),
                         "kernel_speedup_vs_" + args.storage: round(k_avg_ms / k2, 3),
                         "autotuned_tile_order": {k: (round(v, 4) if isinstance(v, float) else v)
                                                  for k, v in tune2.items()}}
            print(f"[bench] {st}: {1e3 * dt2 / args.steps:.3f} ms/step, k_hydro {k2:.3f} ms, tiles {tune2}",
                  file=sys.stderr, flush=True)
        qerr = qoi_error_vs_f64(("f32", "f16", "bf16"), args.collision, Q)
        if qerr is not None:
            sweep["qoi_vs_f64_oracle"] = qerr

    cpu = None
    if f_dev is not None:
        v, secs, info = cpu_baseline(cfg, f_dev, steps_done, args.cpu_steps, args.collision, Q)
        del f_dev
        cpu = {"value": round(v, 3), "unit": "MLUPS", "cores": info["threads_used"], "kind": "port",
               "sample": f"oracle D3Q{Q} {args.collision} fp32 on the same {args.size}^3 tank (IBM + "
                         f"collide-stream), {args.cpu_steps} steps timed after one untimed step, started from "
                         f"the GPU run's populations after {steps_done} steps; {secs:.1f}s",
               **info}

    out = {
        "metric": f"MLUPS & % HBM roofline, D3Q{Q} tank {args.size}^3",
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
                   "fluid_cells": int(fluid), "parallelism": "single GPU",
                   "autotuned_tile_order": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tune.items()}},
        "mlups_total": round(mlups_total, 2),
        "mlups_fluid": round(mlups_fluid, 2),
        "graph_span": {"ms_per_step": round(1e3 * dt_graph / args.steps, 4),
                       "mlups_fluid": round(fluid * args.steps / dt_graph / 1e6, 2)},
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
