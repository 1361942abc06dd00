#!/usr/bin/env python3
"""Writes tests/golden/c2_digest.json: the CPU oracle's run of the 128^3 FP64
static-dust configuration, for the GPU parity test at its full size
(tests/test_gpu_c2.py).

Configuration: 128^3, order 4, FP64, static dust, P(k) ~ k^-2 on
1 <= |k| <= N/4 with rms 1e-3 and seed 9 (cosmograph_amd.ics.static_dust_ics,
the same recipe the test regenerates), dt = 0.1 dx, 200 RK4 steps.

Recorded:
  * setConstraintCalcs (H and M, all 8 statistics) at the output point of
    steps 0, 10, ..., 200 (stepInit + sources, as the reference's output step
    sees the state, sims/static.cc:74-129);
  * after 200 steps, for each of the 18 _p registers: the sum, L2 norm and
    largest magnitude, and the values at 64 seeded probe points.
The oracle is OpenMP C (oracle/liboracle.so); this takes several minutes.
    python tests/golden/make_c2_digest.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
import pyoracle as po  # noqa: E402
from cosmograph_amd.ics import static_dust_ics  # noqa: E402

FIELDS = ["DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23",
          "DIFFgamma33", "DIFFphi", "A11", "A12", "A13", "A22", "A23", "A33", "DIFFK",
          "Gamma1", "Gamma2", "Gamma3", "DIFFalpha"]
N, ORDER, STEPS, EVERY = 128, 4, 200, 10
SPECTRAL_INDEX, RMS, SEED, PROBE_SEED = -2.0, 1e-3, 9, 2024


def probes(n, count=64, seed=PROBE_SEED):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, size=(count, 3)).tolist()


def main():
    t0 = time.time()
    fields, frw = static_dust_ics(N, SPECTRAL_INDEX, RMS, SEED)
    s = po.Sim(po.make_params(N, order=ORDER))
    for name, arr in fields.items():
        s.field(name)[:] = arr
    s.frw_set(frw["phi"], frw["K"])
    for rho, w in frw["fluids"]:
        s.frw_add_fluid(rho, w)
    L = po.lib()
    cons = []
    for step in range(0, STEPS + 1, EVERY):
        if step:
            s.run_steps(EVERY)
        L.or_bssn_step_init(s.h)
        L.or_bssn_clear_src(s.h)
        L.or_static_add_src(s.h)
        L.or_lambda_add_src(s.h)
        c = s.constraints()
        cons.append({"step": step, "H": [float.hex(float(x)) for x in c["H"]],
                     "M": [float.hex(float(x)) for x in c["M"]]})
        print(f"step {step}: H_L2 {c['H'][7]:.6e} ({time.time() - t0:.0f} s)", flush=True)
    pts = probes(N)
    digest = {}
    for f in FIELDS:
        a = s.field(f + "_p")
        digest[f] = {"sum": float.hex(float(a.sum())), "l2": float.hex(float(np.sqrt(np.sum(a * a)))),
                     "max_abs": float.hex(float(np.abs(a).max())),
                     "probes": [float.hex(float(a[i, j, k])) for i, j, k in pts]}
    out = {"config": {"n": N, "order": ORDER, "steps": STEPS, "spectral_index": SPECTRAL_INDEX,
                      "rms": RMS, "seed": SEED, "every": EVERY},
           "probe_points": pts, "constraints": cons, "fields": digest,
           "frw": [float.hex(float(x)) for x in s.frw_get()]}
    with open(os.path.join(HERE, "c2_digest.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"wrote c2_digest.json in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
