"""Restatement of the reference's Bardeen / SVT diagnostics
(components/bssn/bardeen.cc:1-472, bardeen.h:16-236) in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
path's cg_bardeen.  The product package never imports it.

Inputs are the BSSN state as the reference's output step sees it (the _a
register after stepInit and the sources, sims/sim.cc:214-231) and the
right-hand sides ev_* of every field at that state (the _c register after
BSSN::RKEvolve, which the C oracle computes), so the reference's
per-point calls to set_bd_values / ev_* are the same numbers.

Stencils follow utils/math.h (first derivatives :74-586, pure second
derivatives :873-1085, diagonal-only mixed :676-855) in the association
order of the C oracle (oracle/oracle.c), so they agree with it bit for bit
(tests/test_oracle_bardeen.py).  Fourier::inverseLaplacian
(utils/Fourier.h:252-300) divides each mode by -(2 pi |p| / H_LEN_FRAC)^2
with integer lattice momenta p and zeroes the mean; numpy's FFT stands in
for FFTW's (FFTW is not part of this image).

Reference behaviour kept on purpose (they change the outputs):
  * h_0i is overwritten by its time derivative DT_h0I and dt_h_0i is never
    set, so it stays zero (bardeen.cc:198-208);
  * E uses the conformal metric in the beta^2 term (bardeen.cc:213-215);
  * the violation ratios 8-13 always use the matter+Lambda scale factor
    (bardeen.cc:445-452), and 12 mixes the conformal <alpha> with the
    coordinate average of DIFFK;
  * the violation "max" is the largest signed value (math.h:1489-1500);
  * standard deviations divide by N - 1 (math.h:1424-1435).

Parity unpinned: which statistic (average, maximum, standard deviation) and
which normalisation each of the 17 violations takes (bardeen.cc:382-472).
The fields they reduce are pinned by pure scalar, vector and tensor modes,
the matter + Lambda scale factor by its closed form, the second time
derivatives by the trajectory, and lin_viol by lin_viol = 2 (Phi - Psi).
"""
from __future__ import annotations

import math

import numpy as np

PI = 4.0 * math.atan(1.0)  # cosmo_macros.h:107
SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
NUM_VIOLS = 17             # bardeen.h:14


# ------------------------------------------------------------- stencils --
def _P(f, axis, t):
    """f at offset t along axis (periodic): f[i + t]."""
    return np.roll(f, -t, axis=axis)


def _Q(f, a, b, t1, t2):
    return np.roll(np.roll(f, -t1, axis=a), -t2, axis=b)


def d1(f, axis, order, dx):
    """derivative (math.h:1180-1200)."""
    P = lambda t: _P(f, axis, t)  # noqa: E731
    if order == 2:
        return (-1.0 / 2.0 * P(-1) + 1.0 / 2.0 * P(1)) / dx
    if order == 4:
        return (+1.0 / 12.0 * P(-2) - 2.0 / 3.0 * P(-1) + 2.0 / 3.0 * P(1) - 1.0 / 12.0 * P(2)) / dx
    if order == 6:
        return (-1.0 / 60.0 * P(-3) + 3.0 / 20.0 * P(-2) - 3.0 / 4.0 * P(-1) + 3.0 / 4.0 * P(1) -
                3.0 / 20.0 * P(2) + 1.0 / 60.0 * P(3)) / dx
    return ((1.0 / 280.0 * P(-4) - 1.0 / 280.0 * P(4)) - (4.0 / 105.0 * P(-3) - 4.0 / 105.0 * P(3)) +
            (1.0 / 5.0 * P(-2) - 1.0 / 5.0 * P(2)) - (4.0 / 5.0 * P(-1) - 4.0 / 5.0 * P(1))) / dx


def dd(f, a1, a2, order, dx):
    """double_derivative (math.h:1319-1330): pure or diagonal mixed."""
    if a1 == a2:
        P = lambda t: _P(f, a1, t)  # noqa: E731
        if order == 2:
            return (P(-1) - 2.0 * P(0) + P(1)) / dx / dx
        if order == 4:
            return (-1.0 / 12.0 * P(-2) + 4.0 / 3.0 * P(-1) - 5.0 / 2.0 * P(0) + 4.0 / 3.0 * P(1) -
                    1.0 / 12.0 * P(2)) / dx / dx
        if order == 6:
            return (1.0 / 90.0 * P(-3) - 3.0 / 20.0 * P(-2) + 3.0 / 2.0 * P(-1) - 49.0 / 18.0 * P(0) +
                    3.0 / 2.0 * P(1) - 3.0 / 20.0 * P(2) + 1.0 / 90.0 * P(3)) / dx / dx
        return (-1.0 / 560.0 * P(-4) + 8.0 / 315.0 * P(-3) - 1.0 / 5.0 * P(-2) + 8.0 / 5.0 * P(-1) -
                205.0 / 72.0 * P(0) + 8.0 / 5.0 * P(1) - 1.0 / 5.0 * P(2) + 8.0 / 315.0 * P(3) -
                1.0 / 560.0 * P(4)) / dx / dx
    a, b = min(a1, a2), max(a1, a2)

    def Q(m):
        return (-_Q(f, a, b, m, -m) + _Q(f, a, b, m, m) + _Q(f, a, b, -m, -m) - _Q(f, a, b, -m, m))
    if order == 2:
        return (Q(1)) / 4.0 / dx / dx
    if order == 4:
        return ((Q(1)) - 1.0 / 16.0 * (Q(2))) / 3.0 / dx / dx
    if order == 6:
        return (135.0 * (Q(1)) - 27.0 / 2.0 * (Q(2)) + (Q(3))) / 360.0 / dx / dx
    return (2.0 / 5.0 * (Q(1)) - 1.0 / 20.0 * (Q(2)) + 2.0 / 315.0 * (Q(3)) - 1.0 / 2240.0 * (Q(4))) / dx / dx


def inverse_laplacian(f, box):
    """Fourier::inverseLaplacian (utils/Fourier.h:252-300): the same momenta
    2 pi p / H_LEN_FRAC along every axis."""
    nx, ny, nz = f.shape
    F = np.fft.rfftn(f)
    px = np.where(np.arange(nx) <= nx // 2, np.arange(nx), np.arange(nx) - nx).astype(np.float64)
    py = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny).astype(np.float64)
    pz = np.arange(nz // 2 + 1).astype(np.float64)
    PX, PY, PZ = np.meshgrid(px, py, pz, indexing="ij")
    pmag = np.sqrt(PX * PX + PY * PY + PZ * PZ) * 2.0 * PI / box
    with np.errstate(divide="ignore", invalid="ignore"):
        F = F / (-pmag * pmag)
    F[0, 0, 0] = 0.0
    return np.fft.irfftn(F, s=f.shape, axes=(0, 1, 2))


# ------------------------------------------------------------ reductions --
def average(f):
    return float(np.sum(f)) / f.size


def conformal_average(f, dphi, phiF=0.0):
    """math.h:1382-1415."""
    w = np.exp(6.0 * (dphi + phiF))
    vol = float(np.sum(w)) / f.size
    return float(np.sum(w * f)) / f.size / vol


def standard_deviation(f, avg):
    return math.sqrt(float(np.sum((avg - f) ** 2)) / (f.size - 1))


def vmax(f):
    return float(np.max(f))


# ------------------------------------------------- FLRW matter + Lambda --
def ml_initial_time(OL):
    if OL == 0:
        return 2.0 / 3.0
    s = math.sqrt(OL)
    return math.log1p(2.0 / (1.0 - s) * s) / (3.0 * s)


def ml_scale_factor(t, OL):
    """bardeen.h:198-213."""
    if OL == 0:
        tI = ml_initial_time(OL)
        return ((tI + t) / tI) ** (2.0 / 3.0)
    s = math.sqrt(OL)
    return (math.cosh(3.0 * s * t / 2.0) + math.sinh(3.0 * s * t / 2.0) / s) ** (2.0 / 3.0)


def ml_hubble(t, OL):
    a = ml_scale_factor(t, OL)
    return math.sqrt((1 - OL) * a ** -3.0 + OL)


def ml_d2adt2(t, OL):
    a = ml_scale_factor(t, OL)
    H = ml_hubble(t, OL)
    return a * H * H - (1.0 - OL) * 3.0 / 2.0 / a / a


# ----------------------------------------------------------- potentials --
def set_potentials(a_reg, c_reg, t, order, dx, box, use_ml=True, omega_L=0.0):
    """Bardeen::setPotentials (bardeen.cc:22-460).

    a_reg: the state (name -> array): DIFFgamma_ij, DIFFphi, A_ij, DIFFK,
    DIFFalpha and, with the shift, beta1..3.  c_reg: the ev_* right-hand
    sides of the same names (ev_beta^i with the shift).  Returns
    (fields, viols): the arrays the reference adds to the field map as
    "Bardeen_<name>" and the 17 SVT violation values."""
    D1 = lambda f, d: d1(f, d, order, dx)  # noqa: E731
    DD = lambda f, p, q: dd(f, p, q, order, dx)  # noqa: E731
    shape = a_reg["DIFFphi"].shape
    n = float(np.prod(shape))
    zero = np.zeros(shape)
    dphi_a = a_reg["DIFFphi"]
    beta = [a_reg.get(f"beta{m + 1}", zero) for m in range(3)]
    dt_beta = [c_reg.get(f"beta{m + 1}", zero) for m in range(3)]

    alpha_avg = 1.0 + conformal_average(a_reg["DIFFalpha"], dphi_a)
    K_avg = conformal_average(a_reg["DIFFK"], dphi_a)
    phi_avg = conformal_average(dphi_a, dphi_a)
    if use_ml:
        a = ml_scale_factor(t, omega_L)
        H = ml_hubble(t, omega_L)
        dadt = H * a
        d2adt2 = ml_d2adt2(t, omega_L)
    else:
        a = math.exp(2.0 * phi_avg)
        dadt = -1.0 / 3.0 * a * alpha_avg * K_avg
        H = dadt / a
        d2adt2 = None

    g = {}   # full conformal metric gamma_IJ (symmetric access)
    for I, J in SYM:
        g[(I, J)] = g[(J, I)] = a_reg[f"DIFFgamma{I + 1}{J + 1}"] + (1.0 if I == J else 0.0)
    dt_g = {}
    for I, J in SYM:
        dt_g[(I, J)] = dt_g[(J, I)] = c_reg[f"DIFFgamma{I + 1}{J + 1}"]
    alpha = 1.0 + a_reg["DIFFalpha"]
    K = a_reg["DIFFK"]               # K_FRW = 0 (checked by the caller)
    phi = dphi_a
    dt_phi = c_reg["DIFFphi"]
    dtalpha = c_reg["DIFFalpha"]
    dbeta = [[D1(beta[m], j) for m in range(3)] for j in range(3)]     # [j][m] = d_j beta^m
    ddtbeta = [[D1(dt_beta[m], j) for m in range(3)] for j in range(3)]
    dkbetak = dbeta[0][0] + dbeta[1][1] + dbeta[2][2]
    dkdtbetak = D1(dt_beta[0], 0) + D1(dt_beta[1], 1) + D1(dt_beta[2], 2)

    # d^2/dt^2 gamma~_IJ (bardeen.cc:110-127)
    d2t_g = {}
    for I, J in SYM:
        dgIJ = [D1(a_reg[f"DIFFgamma{I + 1}{J + 1}"], k) for k in range(3)]
        ddtg = [D1(dt_g[(I, J)], k) for k in range(3)]
        v = (-2.0 * dtalpha * a_reg[f"A{I + 1}{J + 1}"] - 2.0 * alpha * c_reg[f"A{I + 1}{J + 1}"]
             + dt_beta[0] * dgIJ[0] + dt_beta[1] * dgIJ[1] + dt_beta[2] * dgIJ[2]
             + beta[0] * ddtg[0] + beta[1] * ddtg[1] + beta[2] * ddtg[2]
             + dt_g[(I, 0)] * dbeta[J][0] + dt_g[(I, 1)] * dbeta[J][1] + dt_g[(I, 2)] * dbeta[J][2]
             + dt_g[(J, 0)] * dbeta[I][0] + dt_g[(J, 1)] * dbeta[I][1] + dt_g[(J, 2)] * dbeta[I][2]
             + g[(I, 0)] * ddtbeta[J][0] + g[(I, 1)] * ddtbeta[J][1] + g[(I, 2)] * ddtbeta[J][2]
             + g[(J, 0)] * ddtbeta[I][0] + g[(J, 1)] * ddtbeta[I][1] + g[(J, 2)] * ddtbeta[I][2]
             - 2.0 / 3.0 * (dt_g[(I, J)] * dkbetak + g[(I, J)] * dkdtbetak))
        d2t_g[(I, J)] = d2t_g[(J, I)] = v
    # d^2/dt^2 phi (bardeen.cc:137-143)
    dphi = [D1(phi, k) for k in range(3)]
    ddtphi = [D1(dt_phi, k) for k in range(3)]
    d2t_phi = (-1.0 / 6.0 * (dtalpha * K + alpha * c_reg["DIFFK"])
               + dt_beta[0] * dphi[0] + dt_beta[1] * dphi[1] + dt_beta[2] * dphi[2]
               + beta[0] * ddtphi[0] + beta[1] * ddtphi[1] + beta[2] * ddtphi[2]
               + 1.0 / 6.0 * dkdtbetak)
    if not use_ml:
        d2adt2 = H * dadt + 2.0 * a * conformal_average(d2t_phi, dphi_a)

    # h_ij, time derivatives, h_0i, E (bardeen.cc:151-216)
    e4phi = np.exp(4.0 * phi)
    h, dt_h, d2t_h = {}, {}, {}
    for I, J in SYM:
        diag = I == J
        h[(I, J)] = e4phi * g[(I, J)] - (a * a if diag else 0.0)
        dt_h[(I, J)] = e4phi * (4.0 * dt_phi * g[(I, J)] + dt_g[(I, J)]) - (2.0 * a * dadt if diag else 0.0)
        d2t_h[(I, J)] = (e4phi * ((4.0 * dt_phi) ** 2 * g[(I, J)] + 8.0 * dt_phi * dt_g[(I, J)] +
                                  4.0 * d2t_phi * g[(I, J)] + d2t_g[(I, J)])
                         - (2.0 * (dadt * dadt + a * d2adt2) if diag else 0.0))
        for T in (h, dt_h, d2t_h):
            T[(J, I)] = T[(I, J)]
    h0 = []
    for I in range(3):
        h0I = e4phi * (g[(I, 0)] * beta[0] + g[(I, 1)] * beta[1] + g[(I, 2)] * beta[2])
        h0.append(4.0 * dt_phi * h0I + e4phi * (dt_g[(I, 0)] * beta[0] + dt_g[(I, 1)] * beta[1] +
                                                dt_g[(I, 2)] * beta[2] + g[(I, 0)] * dt_beta[0] +
                                                g[(I, 1)] * dt_beta[1] + g[(I, 2)] * dt_beta[2]))
    dt_h0 = [zero, zero, zero]
    E = (alpha * alpha - 1.0 - g[(0, 0)] * beta[0] * beta[0] - g[(1, 1)] * beta[1] * beta[1]
         - g[(2, 2)] * beta[2] * beta[2]
         - 2.0 * (g[(0, 1)] * beta[0] * beta[1] + g[(0, 2)] * beta[0] * beta[2] + g[(1, 2)] * beta[1] * beta[2]))

    # A (bardeen.cc:220-287)
    def ddsum(T):
        return (DD(T[(0, 0)], 0, 0) + DD(T[(1, 1)], 1, 1) + DD(T[(2, 2)], 2, 2)
                + 2.0 * (DD(T[(0, 1)], 0, 1) + DD(T[(0, 2)], 0, 2) + DD(T[(1, 2)], 1, 2)))
    A_, dt_A, d2t_A = (inverse_laplacian(ddsum(T), box) for T in (h, dt_h, d2t_h))
    h_tr = h[(0, 0)] + h[(1, 1)] + h[(2, 2)]
    dt_h_tr = dt_h[(0, 0)] + dt_h[(1, 1)] + dt_h[(2, 2)]
    d2t_h_tr = d2t_h[(0, 0)] + d2t_h[(1, 1)] + d2t_h[(2, 2)]
    A_ = (h_tr - A_) / a / a / 2.0
    dt_A = -2.0 * A_ * H + (dt_h_tr - dt_A) / a / a / 2.0
    d2t_A = -2.0 * A_ * (H ** 2 + d2adt2 / a) - 4.0 * H * dt_A + (d2t_h_tr - d2t_A) / a / a / 2.0
    avg_A, avg_dt_A, avg_d2t_A = average(A_), average(dt_A), average(d2t_A)
    avg_h_tr, avg_dt_h_tr, avg_d2t_h_tr = float(np.sum(h_tr)) / n, float(np.sum(dt_h_tr)) / n, \
        float(np.sum(d2t_h_tr)) / n
    A_ = A_ + (avg_h_tr / 3.0 / a / a - avg_A)
    dt_A = dt_A + (avg_dt_h_tr / 3.0 / a / a - 2.0 * avg_h_tr * dadt / 3.0 / a / a / a - avg_dt_A)
    d2t_A = d2t_A + (avg_d2t_h_tr / 3.0 / a / a - 2.0 * dadt * avg_dt_h_tr / 3.0 / a / a / a
                     + 6.0 * avg_h_tr * dadt * dadt / 3.0 / a / a / a / a - 2.0 * avg_h_tr * d2adt2 / 3.0 / a / a / a
                     - 2.0 * avg_dt_h_tr * dadt / 3.0 / a / a / a - avg_d2t_A)

    # B (bardeen.cc:290-309)
    B = inverse_laplacian(h_tr / a / a - 3.0 * A_, box)
    dt_B = inverse_laplacian(dt_h_tr / a / a - 2.0 / a / a / a * dadt * h_tr - 3.0 * dt_A, box)
    d2t_B = inverse_laplacian(d2t_h_tr / a / a - 4.0 / a / a / a * dadt * dt_h_tr - 2.0 / a / a / a * d2adt2 * h_tr
                              + 6.0 / a / a / a / a * dadt * dadt * h_tr - 3.0 * d2t_A, box)
    # F (bardeen.cc:312-325)
    F_raw = (D1(h0[0], 0) + D1(h0[1], 1) + D1(h0[2], 2)) / a
    dt_F = inverse_laplacian((D1(dt_h0[0], 0) + D1(dt_h0[1], 1) + D1(dt_h0[2], 2)) / a - H * F_raw, box)
    F = inverse_laplacian(F_raw, box)
    # Bardeen potentials (bardeen.cc:329-340)
    Psi = -1.0 / 2.0 * (A_ + 2.0 * H * (a * F - 0.5 * a * a * dt_B))
    Phi = 1.0 / 2.0 * (E + 2.0 * (dadt * F + a * dt_F) - 2.0 * a * dadt * dt_B - a * a * d2t_B)

    # vector potentials (bardeen.cc:344-392)
    G = [D1(F, I) - h0[I] / a for I in range(3)]
    dA = [D1(A_, I) for I in range(3)]
    ddtA = [D1(dt_A, I) for I in range(3)]
    others = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
    C, dt_C = [], []
    for I in range(3):
        p, q = others[I]
        c_raw = (D1(h[(I, p)], p) + D1(h[(I, q)], q) - D1(h[(p, p)], I) - D1(h[(q, q)], I)) / a / a + 2.0 * dA[I]
        dtc_raw = ((D1(dt_h[(I, p)], p) + D1(dt_h[(I, q)], q) - D1(dt_h[(p, p)], I) - D1(dt_h[(q, q)], I)) / a / a
                   - 2.0 * H * (c_raw - 2.0 * dA[I]) + 2.0 * ddtA[I])
        C.append(c_raw)
        dt_C.append(dtc_raw)
    C = [inverse_laplacian(x, box) for x in C]
    dt_C = [inverse_laplacian(x, box) for x in dt_C]
    Vmag = np.sqrt((G[0] - a * dt_C[0]) ** 2 + (G[1] - a * dt_C[1]) ** 2 + (G[2] - a * dt_C[2]) ** 2)
    # tensor potentials (bardeen.cc:395-409)
    Dt = {}
    for I, J in SYM:
        Dt[(I, J)] = (h[(I, J)] / a / a - (1.0 if I == J else 0.0) * A_ - DD(B, I, J)
                      - D1(C[J], I) - D1(C[I], J))

    # linearised constraint violation (bardeen.cc:412-455)
    def abs_dder(f):
        return (np.abs(DD(f, 0, 0)) + np.abs(DD(f, 0, 1)) + np.abs(DD(f, 0, 2)) + np.abs(DD(f, 1, 1))
                + np.abs(DD(f, 1, 2)) + np.abs(DD(f, 2, 2)))
    lin_viol = E + A_ - a * a * d2t_B - 3.0 * a * dadt * dt_B + 2.0 * a * dt_F + 4.0 * dadt * F
    lin_viol_mag = np.sqrt(E ** 2 + A_ ** 2 + (a * a * d2t_B) ** 2 + (3.0 * a * dadt * dt_B) ** 2
                           + (2.0 * a * dt_F) ** 2 + (4.0 * dadt * F) ** 2)
    lin_viol_der_mag = (abs_dder(E) + abs_dder(A_) + a * a * abs_dder(d2t_B) + abs(3.0 * a * dadt) * abs_dder(dt_B)
                        + 2.0 * a * abs_dder(dt_F) + abs(4.0 * dadt) * abs_dder(F))
    lin_viol_der = abs_dder(lin_viol)

    mean_viol_der = average(lin_viol_der)
    viol_scale_der = average(lin_viol_der_mag)
    max_viol_der = vmax(lin_viol_der)
    std_viol_der = standard_deviation(lin_viol_der, mean_viol_der)
    mean_viol = average(lin_viol)
    viol_scale = average(lin_viol_mag)
    max_viol = vmax(lin_viol)
    std_viol = standard_deviation(lin_viol, mean_viol)
    mlA, mlH, mld2 = ml_scale_factor(t, omega_L), ml_hubble(t, omega_L), ml_d2adt2(t, omega_L)
    q = lambda x, y: float(np.divide(np.float64(x), np.float64(y)))  # noqa: E731  (IEEE: x/0 = +-inf)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = [q(mean_viol_der, viol_scale_der), q(std_viol_der, viol_scale_der), q(max_viol_der, viol_scale_der),
             max_viol_der, q(mean_viol, viol_scale), q(std_viol, viol_scale), q(max_viol, viol_scale), max_viol,
             1.0 - q(mlA, math.exp(2.0 * phi_avg)),
             1.0 - q(mlH, -1.0 / 3.0 * a * alpha_avg * K_avg),
             1.0 - q(mld2, H * dadt + 2.0 * a * conformal_average(d2t_phi, dphi_a)),
             1.0 - q(mlA, math.exp(2.0 * average(dphi_a))),
             1.0 - q(mlH, -1.0 / 3.0 * a * alpha_avg * average(a_reg["DIFFK"])),
             1.0 - q(mld2, H * dadt + 2.0 * a * average(d2t_phi))]
    vavg = average(Vmag)
    v += [vavg, standard_deviation(Vmag, vavg), vmax(Vmag)]
    fields = {"Phi": Phi, "Psi": Psi, "A": A_, "dt_A": dt_A, "d2t_A": d2t_A, "B": B, "dt_B": dt_B, "d2t_B": d2t_B,
              "E": E, "F": F, "G1": G[0], "G2": G[1], "G3": G[2], "C1": C[0], "C2": C[1], "C3": C[2],
              "dt_C1": dt_C[0], "dt_C2": dt_C[1], "dt_C3": dt_C[2], "Vmag": Vmag,
              "D11": Dt[(0, 0)], "D12": Dt[(0, 1)], "D13": Dt[(0, 2)], "D22": Dt[(1, 1)], "D23": Dt[(1, 2)],
              "D33": Dt[(2, 2)]}
    inter = {"d2t_phi": d2t_phi, "h0": h0, "a": a, "dadt": dadt, "H": H, "d2adt2": d2adt2, "lin_viol": lin_viol,
             "lin_viol_der": lin_viol_der}
    for I, J in SYM:
        inter[f"d2t_g{I + 1}{J + 1}"] = d2t_g[(I, J)]
        inter[f"h{I + 1}{J + 1}"] = h[(I, J)]
        inter[f"dt_h{I + 1}{J + 1}"] = dt_h[(I, J)]
        inter[f"d2t_h{I + 1}{J + 1}"] = d2t_h[(I, J)]
    return fields, np.array(v, dtype=np.float64), inter
