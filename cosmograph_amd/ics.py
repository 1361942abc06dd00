"""Constraint-satisfying initial data for BSSN + static dust on a periodic box.

The reference has no solver that takes a density contrast to a metric: its
static ICs go from a chosen phi to rho (components/static/static_ic.cc:185-199,
318-326) and its multigrid solver is an absent submodule.  This generator is
this project's own recipe, shared by the GPU runs and the CPU oracle:

1. delta(x): Gaussian random field, P(k) ~ k^n on 1 <= |k| <= N/4 (|k| in
   units of 2 pi / L), complex normal mode amplitudes drawn from a seeded
   MT19937 stream, rescaled to the requested rms.
2. Conformally flat metric e^{4 phi} delta_ij with K = K_FRW = -sqrt(24 pi
   rho_bar), A_ij = 0, Gamma^i = 0, alpha = 1.  The Hamiltonian constraint
   (bssn.cc:1081-1083) then reads  lap(e^phi) = -2 pi e^{5 phi} drho,
   solved spectrally by fixed-point iteration on u = e^phi with the
   solvability condition  <u^5 drho> = 0  enforced by a uniform shift of
   drho (which keeps the requested delta up to that constant).
3. Static dust: D = e^{6 phi} rho is conserved; DIFFD = D - rho_bar
   (the reference's DIFFr = e^{-6 phi} DIFFD + rho_FRW expm1(-6 phi),
   components/static/static.cc:32-33).
4. FRW reference: phi = 0, K = K_FRW, one w = 0 fluid of density rho_bar
   (components/static/static_ic.cc:372-378).
"""
from __future__ import annotations

import numpy as np

RHO_BAR = 3.0 / (8.0 * np.pi)  # 3 H_I^2 / (8 pi) with H_I = 1


def _kgrid(n, box):
    k1 = np.fft.fftfreq(n, d=1.0 / n)          # integer wavenumbers
    kz = np.fft.rfftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k1, k1, kz, indexing="ij")
    return np.sqrt(KX**2 + KY**2 + KZ**2), (2.0 * np.pi / box) ** 2 * (KX**2 + KY**2 + KZ**2)


def gaussian_modes(n: int, spectral_index: float, seed: int) -> np.ndarray:
    """Half-spectrum [n][n][n/2+1] of the density contrast before scaling:
    |k|^(n_s/2) (g1 + i g2) on the shell 1 <= |k| <= N/4, g from a seeded
    MT19937 stream (the role of the reference's f_field, ICs/ICs.cc:38-134)."""
    kmag, _ = _kgrid(n, 1.0)
    rng = np.random.Generator(np.random.MT19937(seed))
    amp = np.zeros_like(kmag)
    shell = (kmag >= 1.0) & (kmag <= n / 4.0)
    amp[shell] = kmag[shell] ** (spectral_index / 2.0)
    return amp * (rng.standard_normal(kmag.shape) + 1j * rng.standard_normal(kmag.shape))


def gaussian_delta(n: int, spectral_index: float, rms: float, seed: int, box: float = 1.0) -> np.ndarray:
    modes = gaussian_modes(n, spectral_index, seed)
    delta = np.fft.irfftn(modes, s=(n, n, n), axes=(0, 1, 2))
    delta -= delta.mean()
    s = np.sqrt(np.mean(delta**2))
    return delta * (rms / s) if s > 0 else delta


def solve_conformal_factor(drho: np.ndarray, box: float = 1.0, tol: float = 1e-15, max_iter: int = 200):
    """Solve lap(u) = -2 pi u^5 (drho - c), mean(u) = 1, for u and the shift c."""
    n = drho.shape[0]
    _, k2 = _kgrid(n, box)
    inv = np.zeros_like(k2)
    inv[k2 > 0] = -1.0 / k2[k2 > 0]
    u = np.ones_like(drho)
    c = 0.0
    for _ in range(max_iter):
        u5 = u**5
        c = float(np.mean(u5 * drho) / np.mean(u5))
        rhs = -2.0 * np.pi * u5 * (drho - c)
        u_new = 1.0 + np.fft.irfftn(np.fft.rfftn(rhs) * inv, s=drho.shape, axes=(0, 1, 2))
        d = np.max(np.abs(u_new - u))
        u = u_new
        if d < tol:
            break
    return u, c


def static_dust_ics(n: int, spectral_index: float = -2.0, rms: float = 1e-3, seed: int = 9,
                    box: float = 1.0):
    """Returns (fields, frw): fields maps reference keys ("DIFFphi_p", "DIFFD_a", ...)
    to float64 arrays of shape (n, n, n); frw = dict(phi, K, fluids=[(rho, w)])."""
    delta = gaussian_delta(n, spectral_index, rms, seed, box)
    drho = RHO_BAR * delta
    u, c = solve_conformal_factor(drho, box)
    phi = np.log(u)
    rho = RHO_BAR + drho - c
    fields = {"DIFFphi_p": phi, "DIFFD_a": np.exp(6.0 * phi) * rho - RHO_BAR}
    K_frw = -np.sqrt(24.0 * np.pi * RHO_BAR)
    frw = {"phi": 0.0, "K": K_frw, "fluids": [(RHO_BAR, 0.0)]}
    return fields, frw


def static_dust_ics_device(g, spectral_index: float = -2.0, rms: float = 1e-3, seed: int = 9,
                           tol: float = 1e-15, max_iter: int = 200):
    """static_dust_ics computed on the GPU context g (cosmograph_amd.BSSN of an
    n^3 grid): the same modes, then the inverse FFT, the constraint solve and
    the register writes on the device (cg_ics_static_dust).  Returns (c, iterations)."""
    n = g.params.shape()[0]
    return g.ics_static_dust(gaussian_modes(n, spectral_index, seed), rms, tol, max_iter)


def save_ics(directory: str, fields: dict, frw: dict, lam: float = 0.0) -> None:
    """Write initial data in the layout the `cosmo_gpu` driver reads
    (`ic_dir` config key): `fields.txt` lists the field-map names, each
    `<name>.f64` holds its values as raw little-endian doubles in the grid's
    index order, and `frw.txt` holds "phi K", one "fluid rho w" line per
    fluid and "lambda L"."""
    import os
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "fields.txt"), "w") as fh:
        for name, arr in fields.items():
            np.ascontiguousarray(arr, dtype="<f8").tofile(os.path.join(directory, name + ".f64"))
            fh.write(name + "\n")
    with open(os.path.join(directory, "frw.txt"), "w") as fh:
        # repr of a Python float round-trips exactly through strtod
        fh.write(f"{float(frw['phi'])!r} {float(frw['K'])!r}\n")
        for rho, w in frw.get("fluids", []):
            fh.write(f"fluid {float(rho)!r} {float(w)!r}\n")
        fh.write(f"lambda {float(lam)!r}\n")


def main(argv=None):
    """python -m cosmograph_amd.ics OUT_DIR --n 64 [--spectral-index -2] [--rms 1e-3] [--seed 9]"""
    import argparse
    ap = argparse.ArgumentParser(description="static-dust initial data for the cosmo_gpu driver")
    ap.add_argument("out")
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--spectral-index", type=float, default=-2.0)
    ap.add_argument("--rms", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=9)
    ap.add_argument("--box", type=float, default=1.0)
    ap.add_argument("--lambda", dest="lam", type=float, default=0.0)
    a = ap.parse_args(argv)
    fields, frw = static_dust_ics(a.n, a.spectral_index, a.rms, a.seed, a.box)
    save_ics(a.out, fields, frw, a.lam)


if __name__ == "__main__":
    main()
