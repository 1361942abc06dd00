#!/usr/bin/env python3
"""Benchmark: BSSN RK4 point-updates/s at 256^3, FP32 split vs FP64 GPU.

Workload (the 256^3 split-precision configuration): BSSN + static dust on a
periodic 256^3 grid, box L = 1, dx = 1/256, dt = 0.1 dx, 6th-order centred
stencils, synchronous gauge, delta ~ P(k) ~ k^-3 on 1 <= |k| <= N/4 with
rms 1e-2 (seed 9), constraint-solved metric.  A "step" is one whole RK4 step
(stepInit + 4 fused stage kernels) of the whole grid, inputs resident in HBM.

    python bench.py [--steps K] [--warmup W] [--n 256] [--order 6]

Prints ONE JSON line.  `value` is the rate of the headline storage mode
(--precision, default fp32_fix: every perturbation field stored in 4 bytes,
the conformal metric as 32-bit fixed point, which holds every field within
1e-5 of FP64); the FP64 drop-in is timed on the same initial data and
reported beside it (`fp64_value`, `fp32_over_fp64`), with the relative L2
difference of every field between the two after the same number of steps,
and so are the other 4-byte modes (`fp32_split`: FP64 conformal metric;
`fp32_all`: FP32 metric).
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

FIELDS = ["DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23",
          "DIFFgamma33", "DIFFphi", "A11", "A12", "A13", "A22", "A23", "A33", "DIFFK",
          "Gamma1", "Gamma2", "Gamma3", "DIFFalpha"]
HBM_PEAK_GBS = 8000.0
# what each storage mode computes in (the headline's `dtype`)
DTYPES = {
    "fp32_fix": "fp32 perturbation storage, 4 bytes per value: A_ij, K, Gamma^i, alpha, D in FP32; conformal metric "
                "(DIFFgamma_ij, DIFFphi) in 32-bit fixed point (per-field power-of-two scale); FP32-FMA stencils "
                "(metric: of exact integer differences); FP64 RK4 accumulation/FRW/reductions",
    "fp32_split": "fp32 split: f32 storage+stencils of A_ij, K, Gamma^i, alpha, D; FP64 metric "
                  "(DIFFgamma_ij, DIFFphi), FP64 RK4 accumulation/FRW/reductions",
    "fp32_all": "fp32: every perturbation field (metric included) in FP32 storage and FP32-FMA stencils; "
                "FP64 RK4 accumulation/FRW/reductions",
}
F_FIELDS, G_FIELDS = 18, 1  # evolved fields, pointwise matter fields read per stage


def b_alg_words():
    """Algorithmic words per point-update for the per-stage-fused 4-register
    RK4: stage 1 reads p, writes f and a' (3F); stages 2-3 read a, p, f and
    write f, a' (5F each); stage 4 reads a, p, f and writes p (4F); every
    stage reads the G pointwise matter fields."""
    return 17 * F_FIELDS + 4 * G_FIELDS


def make_ctx(cg, n, order, precision, modes, rms):
    """A context with the static-dust ICs of `modes` computed on the device
    (cg_ics_static_dust: inverse FFT, constraint solve, register writes)."""
    g = cg.BSSN(cg.GpuParams(n=n, order=order, precision=precision, matter="static"))
    t0 = time.perf_counter()
    _, iters = g.ics_static_dust(modes, rms)
    print(f"[bench] device ICs ({precision}): {iters} iterations, {time.perf_counter() - t0:.2f} s",
          file=sys.stderr, flush=True)
    return g


def time_steps(g, warmup, steps, autotune=True):
    tune = g.autotune(3) if autotune else None
    g.step(warmup)
    g.sync()
    g.set_timing(True)
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    t1 = time.perf_counter()
    g.set_timing(False)
    avg_ms, nlaunch = g.stage_time_ms()
    return t1 - t0, avg_ms, nlaunch, tune


def cpu_baseline(n_sample, steps, order, spectral_index, rms, seed):
    """The CPU oracle on a bounded sample: `steps` RK4 steps of the same
    physics on an n_sample^3 grid.  Timed with the oracle's -O3 -ffast-math
    build (oracle/liboracle_perf.so, the reference's optimisation flags,
    OpenMP); the reference's own BSSN translation unit cannot be built
    without FFTW, so this is the restatement ("port")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from cosmograph_amd.ics import static_dust_ics
    fields, frw = static_dust_ics(n_sample, spectral_index, rms, seed)
    s = pyoracle.Sim(pyoracle.make_params(n_sample, order=order), perf=True)
    for name, arr in fields.items():
        s.field(name)[:] = arr
    del fields
    s.frw_set(frw["phi"], frw["K"])
    for rho, w in frw["fluids"]:
        s.frw_add_fluid(rho, w)
    t0 = time.perf_counter()
    s.run_steps(steps)
    dt = time.perf_counter() - t0
    threads = int(s.L.or_num_threads())
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"value": n_sample ** 3 * steps / dt, "unit": "point-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} RK4 steps of the oracle (C restatement, -O3 -ffast-math, OpenMP {threads} threads; "
                      f"{ncpu} host CPUs available) on a {n_sample}^3 grid, same physics/order; {dt:.1f} s"}


def copy_probe(gib):
    """One device-to-device copy of a known size through the HIP runtime, so
    that FETCH_SIZE / WRITE_SIZE can be scaled against known bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(gib * (1 << 30))
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_bytes": n}))


ENSEMBLE = [(48, 48), (24, 96), (8, 128)]  # (realisations, grid side)


def ensemble_leg(cg, steps, warmup, log):
    """Batched ensemble: 48 x 48^3 + 24 x 96^3 + 8 x 128^3 = 80 realisations
    (43,319,296 points per step), BSSN + static dust, order 4, FP32 split,
    P(k)~k^-2 with rms drawn uniformly in [1e-4, 1e-2], seeds 1..80; one
    batched launch per RK4 stage, replayed from a HIP graph.  Afterwards the
    same ensemble runs in FP64 (the drop-in path, which the tests hold to the
    oracle at 1e-12) for the per-realisation constraint norms: H and M L2 of
    every realisation in both precisions and their largest relative gap."""
    rng = np.random.default_rng(80)
    params, spec = [], []
    seed = 0
    for count, n in ENSEMBLE:
        for _ in range(count):
            seed += 1
            spec.append((n, float(rng.uniform(1e-4, 1e-2)), seed))
            params.append(cg.GpuParams(n=n, order=4, precision="fp32_split", matter="static"))
    log(f"[bench] ensemble: {len(params)} realisations, ICs")
    b = cg.Batch(params)
    from cosmograph_amd.ics import gaussian_modes
    for m, (n, rms, sd) in zip(b.members, spec):
        m.ics_static_dust(gaussian_modes(n, -2.0, sd), rms)
    tune = b.autotune(3)
    b.step(warmup)
    t0 = time.perf_counter()
    b.step(steps)
    wall = time.perf_counter() - t0
    gpu_ms = b.step_time_ms()
    pts = sum(n ** 3 for n, _, _ in spec)
    nan = sum(m.numNaNs("DIFFphi_p") for m in b.members)
    out = {"workload": "batched ensemble: 48x48^3 + 24x96^3 + 8x128^3, order 4, FP32 split",
           "realisations": len(spec), "points_per_step": pts, "steps": steps,
           "value": pts * steps / wall, "unit": "point-updates/s", "ms_per_step": 1e3 * wall / steps,
           "gpu_ms_per_step": gpu_ms / steps, "kernel": b.kernel_info(), "autotune": tune, "nan_points": nan}
    c32 = b.setConstraintCalcs()
    del b
    # the same realisations in FP64, same number of steps
    log("[bench] ensemble: FP64 run for the per-realisation constraint norms")
    b = cg.Batch([cg.GpuParams(n=n, order=4, precision="fp64", matter="static") for n, _, _ in spec])
    for m, (n, rms, sd) in zip(b.members, spec):
        m.ics_static_dust(gaussian_modes(n, -2.0, sd), rms)
    b.step(warmup)
    b.step(steps)
    c64 = b.setConstraintCalcs()
    del b
    gaps = {}
    for c in ("H", "M"):
        a32, a64 = np.asarray(c32[c])[:, 7], np.asarray(c64[c])[:, 7]
        gaps[c] = float(np.max(np.abs(a32 - a64) / np.abs(a64)))
        out[f"{c}_L2_fp32"] = [float(f"{x:.6e}") for x in a32]
        out[f"{c}_L2_fp64"] = [float(f"{x:.6e}") for x in a64]
    out["constraint_L2_max_rel_gap_fp32_vs_fp64"] = gaps
    log(f"[bench] ensemble: largest relative gap of a realisation's H / M L2, FP32 vs FP64: "
        f"{gaps['H']:.2e} / {gaps['M']:.2e}")
    return out


C5_FIELDS = ["DIFFphi", "DIFFgamma11", "A11", "DIFFK", "Gamma1", "dust_D", "dust_S1"]


def c5_leg(cg, args, log):
    """512^3 BSSN + dust fluid (D, S_i) + Lambda (energy density 0.7 rho_bar,
    i.e. Lambda = 0.7 * 3 H_I^2), 6th-order stencils, KO coefficient 0.1
    (the reference's KO is identically zero at order 6), delta ~ k^-3 with rms
    3e-2, ICs on the device (cg_ics_static_dust: the fluid at rest), evolved
    as perturbations of the FRW fluids (cg_set_dust_background: the mean dust
    and Lambda as a w = -1 fluid in the FP64 background, the perturbations in
    4-byte storage).  The headline storage mode (--precision: the conformal
    metric in 32-bit fixed point by default), the FP64-metric split mode and
    FP64 runs of the same ICs; reports the rates, the stage kernel's HBM
    fraction of B_alg = (17*22)*4 B per point-update, the Hamiltonian-
    constraint growth (H L2 after / before) and the relative L2 of the 4-byte
    runs against FP64.  Prints one JSON line."""
    from cosmograph_amd.ics import RHO_BAR, gaussian_modes
    n, order, steps = args.n, args.order, args.steps
    lam = 0.7 * RHO_BAR
    log(f"[c5] {n}^3 dust + Lambda, order {order}, {steps} steps; modes")
    modes = gaussian_modes(n, args.spectral_index, args.seed)
    head = args.precision
    precs = [head] + [p for p in ("fp32_split",) if p != head] + ["fp64"]
    runs, keep, tunes = {}, {}, {}
    for prec in precs:
        g = cg.BSSN(cg.GpuParams(n=n, order=order, precision=prec, matter="dust", lambda_=lam, ko_coeff=0.1))
        g.set_dust_background(True)   # perturbations of the FRW fluids (mean dust, Lambda as w = -1)
        t0 = time.perf_counter()
        _, iters = g.ics_static_dust(modes, args.rms)
        log(f"[c5] {prec}: device ICs {iters} iterations, {time.perf_counter() - t0:.1f} s")
        h0 = g.setConstraintCalcs()["H"][7]
        t, k_ms, nl, tune = time_steps(g, args.warmup, steps, not args.no_autotune)
        tunes[f"{prec}_autotune"] = tune
        h1 = g.setConstraintCalcs()["H"][7]
        runs[prec] = {"value": n ** 3 * steps / t, "ms_per_step": 1e3 * t / steps, "stage_kernel_ms": k_ms,
                      "H_L2_start": h0, "H_L2_end": h1, "H_growth": h1 / h0 if h0 else None,
                      "kernel": g.kernel_info(), "nan_points": g.numNaNs("dust_D_p")}
        log(f"[c5] {prec}: {runs[prec]['ms_per_step']:.2f} ms/step, H growth {runs[prec]['H_growth']:.4g}")
        if prec != "fp64":
            keep[prec] = {f: g.download(f + "_p") for f in C5_FIELDS}
        else:
            for p4, fields in keep.items():
                errs = {}
                for f in C5_FIELDS:
                    b = g.download(f + "_p")
                    nb = float(np.sqrt(np.sum(b * b)))
                    d = float(np.sqrt(np.sum((fields[f] - b) ** 2)))
                    errs[f] = float(f"{(d / nb if nb > 0 else d):.3e}")
                runs[p4]["rel_l2_vs_fp64"] = errs
        del g
    b_alg = (17 * 22) * 4   # 4-byte words of the dust-fluid step (F = 22, G = 0)
    achieved = b_alg / 4.0 * n ** 3 / (runs[head]["stage_kernel_ms"] * 1e-3) / 1e9
    r32, r64 = runs[head], runs["fp64"]
    line = {
        "metric": f"BSSN + dust fluid + Lambda RK4 point-updates/s at {n}^3, FP32 split",
        "value": r32["value"], "unit": "point-updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPES[head], "data": "synthetic: seeded Gaussian random field ICs, constraint-solved",
        "config": {"workload": f"dust fluid + Lambda {n}^3, order {order}, k^{args.spectral_index}, "
                               f"rms {args.rms}, Lambda energy density 0.7 rho_bar; KO coefficient 0.1, "
                               f"which is zero at order 6 as the reference's KO stencil; fluid D and S_i "
                               f"only (no energy variable E, as the reference)", "precision": head},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "b_alg_bytes_per_point_update": b_alg,
        "fp32": r32, "fp64": r64, "fp32_over_fp64": r32["value"] / r64["value"]}
    for p4 in keep:
        if p4 != head:
            line[p4] = runs[p4]
    emit(line, tunes, args.detail_file, log)


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", choices=["c3", "c5"], default="c3",
                    help="c3: the 256^3 static-dust workload (default); c5: 512^3 dust fluid + Lambda")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 200 for c3, 400 for c5)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=None, help="grid side (default 256 for c3, 512 for c5)")
    ap.add_argument("--order", type=int, default=6)
    ap.add_argument("--seed", type=int, default=9)
    ap.add_argument("--rms", type=float, default=None, help="rms of delta (default 1e-2 for c3, 3e-2 for c5)")
    ap.add_argument("--spectral-index", type=float, default=-3.0)
    ap.add_argument("--cpu-sample", type=int, default=256, help="grid side of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=2, help="RK4 steps of the CPU-baseline sample")
    ap.add_argument("--detail-file", default=None, help="also write the line and the detail tables to this file")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-fp32-all", action="store_true", help="skip the all-FP32 storage leg")
    ap.add_argument("--skip-fp32-fix", action="store_true", help="skip the fixed-point metric leg")
    ap.add_argument("--skip-fp32-split", action="store_true", help="skip the FP64-metric split leg")
    ap.add_argument("--precision", choices=sorted(DTYPES), default="fp32_fix",
                    help="storage mode of the headline value (default: the fixed-point metric, the 4-byte mode "
                         "that holds every field within 1e-5 of FP64)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=100, help="batched-ensemble leg (0 = skip)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a device-to-device copy of this many GiB (PMC byte-counter calibration)")
    args = ap.parse_args()
    c5 = args.config == "c5"
    args.steps = args.steps if args.steps is not None else (400 if c5 else 200)
    args.n = args.n if args.n is not None else (512 if c5 else 256)
    args.rms = args.rms if args.rms is not None else (3e-2 if c5 else 1e-2)

    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return

    if args.gpus != 1:
        print("bench.py: measures one GPU only (--gpus 1); sharding a box over several GPUs "
              "(slab decomposition along i with a halo exchange of r planes per stage) is not built.",
              file=sys.stderr)
        sys.exit(2)

    import cosmograph_amd as cg
    n, order = args.n, args.order
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if c5:
        c5_leg(cg, args, log)
        return
    log(f"[bench] ICs {n}^3 P(k)~k^{args.spectral_index} rms {args.rms} seed {args.seed} (modes on the host, "
        f"transform and constraint solve on the device)")
    from cosmograph_amd.ics import gaussian_modes
    modes = gaussian_modes(n, args.spectral_index, args.seed)

    headline = args.precision
    log(f"[bench] headline: {headline}")
    g32 = make_ctx(cg, n, order, headline, modes, args.rms)
    t32, k32_ms, nl32, tune32 = time_steps(g32, args.warmup, args.steps, not args.no_autotune)
    info32 = g32.kernel_info()
    log(f"[bench] {headline}: {1e3 * t32 / args.steps:.3f} ms/step, kernel {info32}")
    pts = n ** 3
    rate32 = pts * args.steps / t32
    words = b_alg_words()
    bytes_pt_upd = words * 4                       # 4-byte storage of every value
    bytes_per_launch = bytes_pt_upd / 4.0 * pts    # average stage

    res = {}
    detail_extra = {}
    if not args.skip_fp64:
        log("[bench] FP64 drop-in")
        g64 = make_ctx(cg, n, order, "fp64", modes, args.rms)
        t64, k64_ms, _, tune64 = time_steps(g64, args.warmup, args.steps, not args.no_autotune)
        rate64 = pts * args.steps / t64
        f64 = {f: g64.download(f + "_p") for f in FIELDS}
        c64 = g64.setConstraintCalcs()
        detail_extra["fp64_autotune"] = tune64
        del g64

        def rel_errs(g):
            out = {}
            for f in FIELDS:
                a, b = g.download(f + "_p"), f64[f]
                nb = float(np.sqrt(np.sum(b * b)))
                d = float(np.sqrt(np.sum((a - b) ** 2)))
                out[f] = d / nb if nb > 0 else d
            return out
        errs = rel_errs(g32)
        # Hamiltonian / momentum constraint of the final states (L2 = H[7],
        # L-infinity = H[2]); the FP64 run stands for the oracle (held to it at
        # 1e-12 by the tests)
        c32 = g32.setConstraintCalcs()
        cons = {f"{c}_{k}_{p}": float(cc[c][q]) for c in ("H", "M") for k, q in (("L2", 7), ("Linf", 2))
                for p, cc in (("fp32", c32), ("fp64", c64))}
        cons["H_L2_fp32_over_fp64"] = float(c32["H"][7] / c64["H"][7]) if c64["H"][7] else None
        cons["H_Linf_fp32_over_fp64"] = float(c32["H"][2] / c64["H"][2]) if c64["H"][2] else None
        log(f"[bench] H L2 / Linf: {headline} {c32['H'][7]:.6e} / {c32['H'][2]:.6e}, "
            f"FP64 {c64['H'][7]:.6e} / {c64['H'][2]:.6e}")
        res = {"fp64_value": rate64, "fp64_ms_per_step": 1e3 * t64 / args.steps,
               "fp64_stage_kernel_ms": k64_ms, "fp32_over_fp64": rate32 / rate64,
               "max_rel_l2_fp32_vs_fp64": max(errs.values()),
               "rel_l2_fp32_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
               "constraints": cons}
        # the other 4-byte-perturbation modes on the same data: the FP64
        # conformal metric (CG_FP32_SPLIT), the metric in FP32 (CG_FP32_ALL)
        # and the metric in 32-bit fixed point (CG_FP32_FIX)
        for prec, skip in (("fp32_split", args.skip_fp32_split), ("fp32_all", args.skip_fp32_all),
                           ("fp32_fix", args.skip_fp32_fix)):
            if skip or prec == headline:
                continue
            log(f"[bench] {prec} storage")
            ga = make_ctx(cg, n, order, prec, modes, args.rms)
            ta, ka_ms, _, tunea = time_steps(ga, args.warmup, args.steps, not args.no_autotune)
            ea = rel_errs(ga)
            ca = ga.setConstraintCalcs()
            worst = max(ea, key=ea.get)
            res[prec] = {
                "value": pts * args.steps / ta, "ms_per_step": 1e3 * ta / args.steps, "stage_kernel_ms": ka_ms,
                "frac": bytes_per_launch / (ka_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "over_fp64": (pts * args.steps / ta) / rate64,
                "max_rel_l2_vs_fp64": float(f"{ea[worst]:.3e}"), "worst_field": worst,
                "H_L2_over_fp64": float(ca["H"][7] / c64["H"][7]) if c64["H"][7] else None,
                "kernel": ga.kernel_info()}
            log(f"[bench] {prec}: {res[prec]['ms_per_step']:.3f} ms/step, worst field {worst} {ea[worst]:.2e}")
            detail_extra[f"{prec}_autotune"] = tunea
            del ga
        del f64

    achieved = bytes_per_launch / (k32_ms * 1e-3) / 1e9
    traffic = load_traffic(f"bssn_static_dust_{n}_o{order}_{headline}")

    del g32
    ens = None
    if args.ensemble_steps > 0:
        ens = ensemble_leg(cg, args.ensemble_steps, args.warmup, log)
        log(f"[bench] ensemble: {ens['value'] / 1e9:.3f} G point-updates/s")

    cpu = None
    if args.cpu_sample > 0:
        log(f"[bench] CPU baseline sample {args.cpu_sample}^3, {args.cpu_steps} steps")
        cpu = cpu_baseline(args.cpu_sample, args.cpu_steps, order, args.spectral_index, args.rms, args.seed)

    line = {
        "metric": "BSSN RK4 point-updates/s at 256^3, FP32-split vs FP64 GPU at <=1e-5 rel L2 err",
        "value": rate32,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[headline],
        "data": "synthetic: seeded Gaussian random field ICs, constraint-solved",
        "config": {"workload": f"BSSN + static dust, {n}^3 periodic, order {order}, "
                               f"P(k)~k^{args.spectral_index:g}, rms {args.rms:g}, dt=0.1dx, {headline}",
                   "grid": n, "stencil_order": order, "precision": headline},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "b_alg_bytes_per_point_update": bytes_pt_upd,
        "stage_kernel_avg_ms": k32_ms,
        "stage_kernel_launches_timed": nl32,
        "kernel": info32,
    }
    detail = {f"{headline}_autotune": tune32, **detail_extra}
    if res:
        line.update(res)
    if ens is not None:
        detail["ensemble"] = dict(ens)
        line["ensemble"] = {k: ens[k] for k in ("workload", "realisations", "points_per_step", "steps", "value",
                                                 "unit", "ms_per_step", "gpu_ms_per_step", "nan_points",
                                                 "constraint_L2_max_rel_gap_fp32_vs_fp64")}
    emit(line, detail, args.detail_file, log)


def emit(line, detail, detail_file, log):
    """The result line on stdout (compact: the autotune tables and the
    per-realisation arrays go to stderr and, with --detail-file, to a file)."""
    text = json.dumps(line)
    if len(text) > 4096:
        log(f"[bench] warning: result line is {len(text)} bytes")
    log("[bench] detail " + json.dumps(detail))
    if detail_file:
        d = os.path.dirname(os.path.abspath(detail_file))
        os.makedirs(d, exist_ok=True)
        with open(detail_file, "w") as fh:
            json.dump({"line": line, "detail": detail}, fh, indent=1)
    print(text, flush=True)


if __name__ == "__main__":
    main()
