"""Helpers shared by the parity tests: identical initial data for the CPU
oracle and the GPU path, and the comparison metric."""
from __future__ import annotations

import numpy as np

FIELDS = ["DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23",
          "DIFFgamma33", "DIFFphi", "A11", "A12", "A13", "A22", "A23", "A33", "DIFFK",
          "Gamma1", "Gamma2", "Gamma3", "DIFFalpha"]


def smooth_random_state(shape, seed=1, amp=1e-3, kmax=2, with_alpha=True):
    """All 18 BSSN fields (+ DIFFD) as sums of a few random low-k periodic
    modes: every term of the right-hand side is exercised."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(ny)[None, :, None] / ny
    z = np.arange(nz)[None, None, :] / nz
    out = {}
    for f in FIELDS + ["DIFFD"]:
        a = np.zeros(shape)
        for _ in range(4):
            k = rng.integers(-kmax, kmax + 1, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            a += rng.standard_normal() * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + ph)
        out[f] = amp * a
    if not with_alpha:
        out["DIFFalpha"][:] = 0
    return out


def rel_l2(a, b):
    """Relative L2 difference; for an identically-zero reference the absolute
    L2 of a is returned (fields like DIFFalpha under the Static lapse)."""
    nb = np.sqrt(np.sum(b * b))
    d = np.sqrt(np.sum((a - b) ** 2))
    return d / nb if nb > 0 else d


def setup_pair(oracle_mod, gpu_mod, params_kw, state=None, ics=None, gpu_kw=None):
    """Create an oracle Sim and a GPU BSSN with the same parameters and data."""
    po = oracle_mod
    n = params_kw.get("n", 16)
    shape = (n, n, n) if np.isscalar(n) else tuple(n)
    op = po.make_params(shape, order=params_kw.get("order", 4), ko=params_kw.get("ko", 0.0),
                        a_adj=params_kw.get("a_adj", 0.0), k_damp=params_kw.get("k_damp", 0.0),
                        normalize=params_kw.get("normalize", 1), lapse=params_kw.get("lapse", "Static"),
                        matter=params_kw.get("matter", "static"), lam=params_kw.get("lam", 0.0))
    osim = po.Sim(op)
    gp = gpu_mod.GpuParams(n=shape, order=params_kw.get("order", 4),
                           precision=params_kw.get("precision", "fp64"),
                           ko_coeff=params_kw.get("ko", 0.0), a_adj_amp=params_kw.get("a_adj", 0.0),
                           k_damping_amp=params_kw.get("k_damp", 0.0),
                           normalize_metric=params_kw.get("normalize", 1),
                           lapse=params_kw.get("lapse", "Static"),
                           matter=params_kw.get("matter", "static"), lambda_=params_kw.get("lam", 0.0),
                           **(gpu_kw or {}))
    g = gpu_mod.BSSN(gp)
    frw = params_kw.get("frw", {"phi": 0.0, "K": -3.0, "fluids": [(3.0 / (8.0 * np.pi), 0.0)]})
    if ics is not None:
        fields, frw = ics
        for name, arr in fields.items():
            osim.field(name)[:] = arr
            g.upload(name, arr)
    if state is not None:
        for f in FIELDS:
            osim.field(f + "_p")[:] = state[f]
            g.upload(f + "_p", state[f])
        osim.field("DIFFD_a")[:] = state["DIFFD"]
        g.upload("DIFFD_a", state["DIFFD"])
    osim.frw_set(frw["phi"], frw["K"])
    g.frw_set(frw["phi"], frw["K"])
    for rho, w in frw["fluids"]:
        osim.frw_add_fluid(rho, w)
        g.frw_add_fluid(rho, w)
    return osim, g
