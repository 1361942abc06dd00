#!/usr/bin/env python3
"""Writes tests/golden/ref_vectors.json from the reference pieces that build
from their own sources (oracle/_ref, see oracle/Makefile target `ref`).

Vectors (inputs are regenerated from the seeds recorded here):
  * stencils: utils/math.h derivative / double_derivative / KO_dissipation_Q /
    upwind_derivative
    on a seeded random field on the 12x10x14 grid the drivers are compiled
    for, orders 2/4/6/8 -> sha256 of the FP64 output bytes plus 16 sample
    values per array;
  * reductions: average, volume_average, conformal_average, min, max,
    numNaNs, conformal_standard_deviation (utils/math.h:1359-1541);
  * static source: Static::addBSSNSrc (components/static/static.cc:19-36);
  * RK4Register trajectory (utils/RK4Register.h) for x' = c0 cos x + c1 x;
  * FRW<double> stage values (utils/FRW.h) for four fluids.
Run in the container that has the reference tree:
    make -C oracle ref && python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402

SEED_FIELD = 20240917
KO_COEFF = 0.37


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def samples(a):
    flat = np.ascontiguousarray(a).reshape(-1)
    idx = np.linspace(0, flat.size - 1, 16).astype(int)
    return [float.hex(float(flat[i])) for i in idx]


def field_input(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape)


def upwind_speed(shape, order):
    c = field_input(shape, SEED_FIELD + 200 + order)
    c.reshape(-1)[::5] = 0.0
    return np.ascontiguousarray(c)


def main():
    out = {"grid": None, "seed_field": SEED_FIELD, "ko_coeff": KO_COEFF, "stencils": {}}
    for order in (2, 4, 6, 8):
        R = po.ref_lib(order)
        nx, ny, nz, o = po.ref_dims(R)
        out["grid"] = [nx, ny, nz]
        f = field_input((nx, ny, nz), SEED_FIELD + order)
        dx = 1.0 / nx
        R.ref_set_steps(dx, 0.1 * dx)
        d1 = np.zeros((3, nx, ny, nz)); R.ref_derivatives(po.dptr(f), po.dptr(d1))
        d2 = np.zeros((9, nx, ny, nz)); R.ref_double_derivatives(po.dptr(f), po.dptr(d2))
        ko = np.zeros((nx, ny, nz)); R.ref_ko(po.dptr(f), KO_COEFF, po.dptr(ko))
        ent = {"derivative": {"sha256": digest(d1), "samples": samples(d1)},
               "double_derivative": {"sha256": digest(d2), "samples": samples(d2)},
               "ko": {"sha256": digest(ko), "samples": samples(ko)}}
        # upwind_derivative with a signed per-point speed (zero at every 5th point)
        c = upwind_speed((nx, ny, nz), order)
        up = np.zeros((3, nx, ny, nz)); R.ref_upwind(po.dptr(f), po.dptr(c), po.dptr(up))
        ent["upwind"] = {"sha256": digest(up), "samples": samples(up)}
        g = np.ascontiguousarray(0.01 * field_input((nx, ny, nz), SEED_FIELD + 100 + order))
        red = np.zeros(7)
        R.ref_reductions(po.dptr(f), po.dptr(g), 0.0123, po.dptr(red))
        ent["reductions"] = [float.hex(float(x)) for x in red]
        if order == 4:
            D = np.ascontiguousarray(0.05 * field_input((nx, ny, nz), SEED_FIELD + 7))
            src = np.zeros((nx, ny, nz))
            R.ref_static_src(po.dptr(g), po.dptr(D), 0.0123, 0.117, po.dptr(src))
            out["static_src"] = {"phi_frw": 0.0123, "rho_frw": 0.117, "sha256": digest(src),
                                 "samples": samples(src)}
            x0 = np.ascontiguousarray(np.linspace(0.1, 2.0, 7))
            traj = np.zeros((25, 7)); last = np.zeros((4, 7))
            R.ref_rk4(po.dptr(x0), 7, 0.05, 0.8, -0.3, 25, po.dptr(traj), po.dptr(last))
            out["rk4"] = {"x0": x0.tolist(), "h": 0.05, "c0": 0.8, "c1": -0.3, "steps": 25,
                          "traj": [float.hex(float(v)) for v in traj.reshape(-1)],
                          "last_stage_a": [float.hex(float(v)) for v in last.reshape(-1)]}
            rho = np.ascontiguousarray([0.2, 0.1, 0.05, 0.03]); w = np.ascontiguousarray([0.0, 1 / 3, -1.0, -0.5])
            fr = np.zeros((30, 4, 4))
            R.ref_frw(0.0, -3.1, 4, po.dptr(rho), po.dptr(w), 0.01, 30, po.dptr(fr))
            out["frw"] = {"phi0": 0.0, "K0": -3.1, "rho": rho.tolist(), "w": w.tolist(), "h": 0.01,
                          "steps": 30, "stage_values": [float.hex(float(v)) for v in fr.reshape(-1)]}
        out["stencils"][str(order)] = ent
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", os.path.join(HERE, "ref_vectors.json"))


if __name__ == "__main__":
    main()
