"""Reading the oracle's committed digests of the full-size static-dust runs
(tests/golden/<name>_digest.json and <name>_sample.npz, written by
tests/golden/make_digest.py) and comparing a GPU context against them."""
from __future__ import annotations

import json
import os

import numpy as np

from common import FIELDS

HERE = os.path.dirname(os.path.abspath(__file__))


def fhex(h):
    return float.fromhex(h)


def load(name):
    with open(os.path.join(HERE, "golden", f"{name}_digest.json")) as fh:
        d = json.load(fh)
    with np.load(os.path.join(HERE, "golden", f"{name}_sample.npz"), allow_pickle=False) as z:
        sample = {f: z[f] for f in FIELDS}
    return d, sample


_ics_cache = {}


def gpu_context(gpu, d, precision, static_dust_ics):
    """A GPU context holding the digest's initial data (the same seeded
    recipe the oracle ran from)."""
    cfg = d["config"]
    key = (cfg["n"], cfg["spectral_index"], cfg["rms"], cfg["seed"])
    if key not in _ics_cache:
        _ics_cache.clear()
        _ics_cache[key] = static_dust_ics(*key)
    fields, frw = _ics_cache[key]
    g = gpu.BSSN(gpu.GpuParams(n=cfg["n"], order=cfg["order"], precision=precision))
    for name, arr in fields.items():
        g.upload(name, arr)
    g.frw_set(frw["phi"], frw["K"])
    for rho, w in frw["fluids"]:
        g.frw_add_fluid(rho, w)
    return g


def sample_errors(g, d, sample):
    """Per field: relative L2 of (GPU - oracle) on the strided sample (the
    absolute L2 for a field the oracle holds at zero), and the GPU field's
    full L2 norm and largest magnitude against the digest's."""
    s = d["sample_stride"]
    out = {}
    for f in FIELDS:
        a = g.download(f + "_p")
        ref = sample[f]
        sub = a[::s, ::s, ::s]
        nb = float(np.sqrt(np.sum(ref * ref)))
        dd = float(np.sqrt(np.sum((sub - ref) ** 2)))
        l2, mx = fhex(d["fields"][f]["l2"]), fhex(d["fields"][f]["max_abs"])
        out[f] = {"sample": dd / nb if nb > 0 else dd,
                  "l2": abs(float(np.sqrt(np.sum(a * a))) - l2) / l2 if l2 > 0 else float(np.sqrt(np.sum(a * a))),
                  "max": abs(float(np.abs(a).max()) - mx) / mx if mx > 0 else float(np.abs(a).max())}
    return out


def constraint_gap(got, want_hex):
    """Relative gaps of the eight H or M statistics (NaN where both are NaN)."""
    want = [fhex(x) for x in want_hex]
    gaps = []
    for q, w in enumerate(want):
        o = got[q]
        if np.isnan(w):
            gaps.append(0.0 if np.isnan(o) else float("inf"))
        else:
            gaps.append(abs(o - w) / abs(w) if w != 0 else abs(o))
    return gaps, want
