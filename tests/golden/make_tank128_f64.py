"""Generate the FP64 reference series for the storage-precision test.

Runs the oracle (the CPU restatement of the reference path, in f64) on the
128^3 stirred tank at the benchmark's ratios (T = H = 120 lu, 4 baffles
T/10 x T/48, 6-blade Rushton D = C = T/3, tip speed 0.05 lu) with a D3Q7
tracer (Gaussian sigma = 2 lu at 0.8 H on the axis, Sc = 1), and stores the
shaft torque of every step and the tracer CV every 10 steps.
Restatement-generated: the reference ships no such series.

Re_imp = 100 (nu = 6.4e-3 lu): at 128^3 the tank without a turbulence model
diverges for lattice viscosities below about 1e-3 (tip-speed cell Reynolds
number above about 50): the f64 oracle at Re 5000 (nu = 1.3e-4) went
non-finite at step 2498, the GPU at Re 5000 / 2500 / 1250 between steps 2340
and 2840, while Re 625 and below run 17,600 steps.  At Re 100 the tracer
reaches the 95 % mixing criterion after 6.7 revolutions, inside the
reference's band of 3 to 7 (vb_02_passive_scalar_mixing.rs:7-14), so the
series carries both Np and t95.

usage: python tests/golden/make_tank128_f64.py [steps]   (about 0.3 s/step on 8 cores, 88 min;
the committed series: 17,600 steps = 7 impeller revolutions, the upper end
of the reference's mixing-time band, vb_02_passive_scalar_mixing.rs:7-14)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import cases  # noqa: E402

N = 128
EVERY_CV = 10


RE = 100.0


def tank_params():
    T = 120.0
    D = T / 3.0
    omega = 0.05 / (0.5 * D)
    nu = (omega / (2.0 * np.pi)) * D * D / RE
    return nu


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 17600
    # subnormal f64 values (< 2.2e-308: far tracer tails, decayed quiescent
    # deviations) would put x86 on its slow path; flushing them changes no
    # torque or CV digit (checked against an unflushed run of 300 steps)
    cases.O.flush_subnormals(True)
    nu = tank_params()
    tc = cases.tank_case(N, nu=nu, collision="cm", Q=27, sigma=2.0)
    t0 = time.time()
    torque, cv, s = cases.run_tank_oracle(tc, steps, prec=8, every=1, scalar_D=nu, d3q7=True,
                                          cv_every=EVERY_CV, progress=500)
    meta = {"n": N, "steps": steps, "re_imp": RE, "nu": nu, "omega": tc["omega"], "D": tc["D"], "T": tc["T"], "sigma": 2.0,
            "scalar": "d3q7", "cv_every": EVERY_CV, "collision": "cm", "Q": 27, "precision": "f64",
            "seconds": round(time.time() - t0, 1)}
    out = os.path.join(HERE, "tank128_cm_f64.npz")
    np.savez_compressed(out, torque=torque, cv=cv, meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(out, meta)


if __name__ == "__main__":
    main()
