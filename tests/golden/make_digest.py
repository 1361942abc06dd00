#!/usr/bin/env python3
"""Writes the oracle's digests of the two full-size static-dust runs, for the
GPU tests at those sizes (tests/test_gpu_c2.py, tests/test_gpu_c3.py):

  c2: 128^3, order 4, FP64, P(k) ~ k^-2 on 1 <= |k| <= N/4, rms 1e-3,
      seed 9, dt = 0.1 dx, 200 RK4 steps;
  c3: 256^3, order 6, P(k) ~ k^-3 on 1 <= |k| <= N/4, rms 1e-2, seed 9,
      dt = 0.1 dx, 200 RK4 steps (the split-precision workload; the oracle is
      FP64 throughout).

ICs: cosmograph_amd.ics.static_dust_ics (the same recipe the tests
regenerate).  Recorded in tests/golden/<name>_digest.json:
  * setConstraintCalcs (H, M: 8 statistics; G, A, S: 7) at the output point
    of steps 0, 10, ..., 200 (stepInit + sources, as the reference's output
    step sees the state, sims/static.cc:74-129), as float.hex strings;
  * after 200 steps, for each of the 18 _p registers: the sum, L2 norm and
    largest magnitude, and the values at 64 seeded probe points;
  * the FRW background.
and in tests/golden/<name>_sample.npz: every field's _p after 200 steps on
the strided 32^3 sub-lattice a[::s, ::s, ::s], s = N/32 (float64, plain
arrays, numpy.load without pickles).

The oracle is OpenMP C (oracle/liboracle.so); c2 takes ~10 min and c3 about
an hour and a half on 8 cores (~12 GB of memory).
    python tests/golden/make_digest.py c2 [c3]
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
CONFIGS = {
    "c2": {"n": 128, "order": 4, "steps": 200, "spectral_index": -2.0, "rms": 1e-3, "seed": 9, "every": 10},
    "c3": {"n": 256, "order": 6, "steps": 200, "spectral_index": -3.0, "rms": 1e-2, "seed": 9, "every": 10},
}
PROBE_SEED = 2024
SAMPLE_SIDE = 32


def probes(n, count=64, seed=PROBE_SEED):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, size=(count, 3)).tolist()


def run(name):
    cfg = CONFIGS[name]
    n = cfg["n"]
    t0 = time.time()
    fields, frw = static_dust_ics(n, cfg["spectral_index"], cfg["rms"], cfg["seed"])
    s = po.Sim(po.make_params(n, order=cfg["order"]))
    for key, arr in fields.items():
        s.field(key)[:] = arr
    del fields
    s.frw_set(frw["phi"], frw["K"])
    for rho, w in frw["fluids"]:
        s.frw_add_fluid(rho, w)
    L = po.lib()
    cons = []
    for step in range(0, cfg["steps"] + 1, cfg["every"]):
        if step:
            s.run_steps(cfg["every"])
        L.or_bssn_step_init(s.h)
        L.or_bssn_clear_src(s.h)
        L.or_static_add_src(s.h)
        L.or_lambda_add_src(s.h)
        c = s.constraints()
        cons.append({"step": step, **{k: [float.hex(float(x)) for x in c[k]] for k in ("H", "M", "G", "A", "S")}})
        print(f"[{name}] step {step}: H_L2 {c['H'][7]:.6e} ({time.time() - t0:.0f} s)", flush=True)
    pts = probes(n)
    stride = n // SAMPLE_SIDE
    digest, sample = {}, {}
    for f in FIELDS:
        a = s.field(f + "_p")
        digest[f] = {"sum": float.hex(float(a.sum())), "l2": float.hex(float(np.sqrt(np.sum(a * a)))),
                     "max_abs": float.hex(float(np.abs(a).max())),
                     "probes": [float.hex(float(a[i, j, k])) for i, j, k in pts]}
        sample[f] = np.ascontiguousarray(a[::stride, ::stride, ::stride])
    out = {"config": cfg, "sample_stride": stride, "probe_points": pts, "constraints": cons, "fields": digest,
           "frw": [float.hex(float(x)) for x in s.frw_get()]}
    with open(os.path.join(HERE, f"{name}_digest.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    np.savez_compressed(os.path.join(HERE, f"{name}_sample.npz"), **sample)
    print(f"[{name}] wrote {name}_digest.json and {name}_sample.npz in {time.time() - t0:.0f} s", flush=True)


def main():
    names = sys.argv[1:] or ["c2"]
    for name in names:
        if name not in CONFIGS:
            raise SystemExit(f"unknown configuration {name!r} (known: {', '.join(CONFIGS)})")
    for name in names:
        run(name)


if __name__ == "__main__":
    main()
