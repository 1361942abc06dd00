"""Pins the oracle's dust fluid (components/dust_fluid/dust.cc:109-264, the
DustSim sequence sims/dust.cc:53-169) without the reference build (dust.cc
includes bssn.h -> utils/Fourier.h -> fftw3.h, absent here):

* the matter sources are the projections of the dust stress-energy tensor
  T^{mu nu} = rho u^mu u^nu, computed here from first principles;
* the fluid right-hand side on sinusoidal data is the closed form of the
  4th-order stencil symbol acting on the reference's fluxes;
* the momentum source on a curved background (metric-gradient force);
* homogeneous dust is the matter-era background with the reference's source
  lag (W, detg, g_ij of the previous stage), classical RK4 otherwise;
* sum D is conserved to round-off (telescoping central fluxes).
"""
import math

import numpy as np
import pytest

FIELDS = ["DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23",
          "DIFFgamma33", "DIFFphi", "A11", "A12", "A13", "A22", "A23", "A33", "DIFFK",
          "Gamma1", "Gamma2", "Gamma3", "DIFFalpha"]
SY = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}


def sym(i, j):
    return SY[(min(i, j), max(i, j))]


def dust_sim(oracle_mod, n, lam=0.0, order=4):
    s = oracle_mod.Sim(oracle_mod.make_params(n, order=order, matter="dust", lam=lam))
    s.frw_set(0.0, 0.0)  # dust ICs leave the reference FRW empty (dust_ic.cc:23-142)
    return s


def sigma4(k, h):
    """Symbol of the 4th-order first-derivative stencil (math.h:231-263) on
    sin/cos: d/dx sin(kx) -> sigma cos(kx)."""
    return (8.0 * math.sin(k * h) - math.sin(2.0 * k * h)) / (6.0 * h)


def populate_sources(oracle_mod, s):
    L = oracle_mod.lib()
    L.or_bssn_step_init(s.h)
    L.or_dust_step_init(s.h)
    L.or_bssn_clear_src(s.h)
    L.or_dust_add_src(s.h)
    L.or_lambda_add_src(s.h)


def test_sources_are_dust_stress_energy(oracle_mod):
    """E = n.T.n, S_i = -gamma_i.T.n, S_ij = gamma_i.T.gamma_j of
    T = rho u u with rho, u_i from D = sqrt(gamma) rho W, S_i = D u_i
    (the conserved variables the fluid evolves)."""
    n = 8
    s = dust_sim(oracle_mod, n)
    rng = np.random.default_rng(3)
    shape = (n, n, n)
    # a generic point-wise metric (unit determinant conformal part) and fluid
    g = {f: 1e-2 * rng.standard_normal(shape) for f in FIELDS[:6]}
    phi = 0.05 * rng.standard_normal(shape)
    for f in FIELDS[:6]:
        s.field(f + "_p")[:] = g[f]
    s.field("DIFFphi_p")[:] = phi
    D = 1.0 + 0.1 * rng.random(shape)
    S = [0.05 * rng.standard_normal(shape) for _ in range(3)]
    s.field("dust_D_p")[:] = D
    for a in range(3):
        s.field(f"dust_S{a + 1}_p")[:] = S[a]
    populate_sources(oracle_mod, s)
    # the oracle normalises the metric in _a (stepInit); use that metric
    gt = np.zeros(shape + (3, 3))
    for i in range(3):
        for j in range(3):
            gt[..., i, j] = s.field(FIELDS[sym(i, j)] + "_a") + (1.0 if i == j else 0.0)
    gam = np.exp(4.0 * phi)[..., None, None] * gt            # physical 3-metric
    gami = np.linalg.inv(gam)
    sqrtg = np.sqrt(np.linalg.det(gam))
    u_lo = np.stack(S, axis=-1) / D[..., None]                 # u_i = S_i / D
    W = np.sqrt(1.0 + np.einsum("...ij,...i,...j", gami, u_lo, u_lo))
    rho = D / (sqrtg * W)
    # T^{mu nu} = rho u^mu u^nu with alpha = 1, beta = 0: n_mu = (-1, 0),
    # u^0 = W, u^i = gamma^{ij} u_j
    E = rho * W * W
    Si = rho[..., None] * W[..., None] * u_lo
    Sij = rho[..., None, None] * u_lo[..., :, None] * u_lo[..., None, :]
    trS = np.einsum("...ij,...ij", gami, Sij)
    np.testing.assert_allclose(s.field("DIFFr_a"), E, rtol=1e-13)
    # S = rho (W^2 - 1) cancels where u is small: compare on the scale of E
    np.testing.assert_allclose(s.field("DIFFS_a"), trS, rtol=1e-12, atol=1e-13 * E.max())
    for a in range(3):
        np.testing.assert_allclose(s.field(f"S{a + 1}_a"), Si[..., a], rtol=1e-12, atol=1e-18)
    for i in range(3):
        for j in range(i, 3):
            tf = Sij[..., i, j] - gam[..., i, j] * trS / 3.0
            np.testing.assert_allclose(s.field(f"STF{i + 1}{j + 1}_a"), tf, rtol=1e-10, atol=1e-13 * E.max())


def test_fluid_rhs_closed_form_flat(oracle_mod):
    """Flat metric, D = D0 + d cos(kx), S1 = e sin(kx): the fluxes
    alpha D v^1 and alpha S1 v^1 are trigonometric polynomials, so the
    stencil result is exact in terms of sigma(k), sigma(2k)."""
    n, m = 16, 2
    s = dust_sim(oracle_mod, n)
    h = 1.0 / n
    k = 2 * math.pi * m
    x = (np.arange(n) * h)[:, None, None] * np.ones((n, n, n))
    D0, d, e = 1.3, 0.2, 0.07
    s.field("dust_D_p")[:] = D0 + d * np.cos(k * x)
    s.field("dust_S1_p")[:] = e * np.sin(k * x)
    L = oracle_mod.lib()
    L.or_bssn_step_init(s.h)
    L.or_dust_step_init(s.h)
    L.or_dust_rk_evolve(s.h)
    # dust.cc:203-224: dt D = sum_i d_i(alpha D v^i), v^1 = S1 on a flat metric
    want_D = D0 * e * sigma4(k, h) * np.cos(k * x) + d * e / 2 * sigma4(2 * k, h) * np.cos(2 * k * x)
    want_S1 = e * e / 2 * sigma4(2 * k, h) * np.sin(2 * k * x)
    np.testing.assert_allclose(s.field("dust_D_c"), want_D, rtol=0, atol=1e-13)
    np.testing.assert_allclose(s.field("dust_S1_c"), want_S1, rtol=0, atol=1e-13)
    assert np.abs(s.field("dust_S2_c")).max() == 0 and np.abs(s.field("dust_S3_c")).max() == 0


def test_momentum_source_on_conformal_factor(oracle_mod):
    """Flat conformal metric, phi = p cos(kx), uniform D, S1: the momentum
    RHS is d_1(alpha S1 v^1) plus the metric-gradient force
    (1/2) alpha W D e^{4 phi} v^i v^j (4 gamma_ij d_1 phi + d_1 gamma_ij)."""
    n, m = 16, 1
    s = dust_sim(oracle_mod, n)
    h = 1.0 / n
    k = 2 * math.pi * m
    x = (np.arange(n) * h)[:, None, None] * np.ones((n, n, n))
    p, D0, s0 = 0.01, 1.1, 0.05
    phi = p * np.cos(k * x)
    s.field("DIFFphi_p")[:] = phi
    s.field("dust_D_p")[:] = D0
    s.field("dust_S1_p")[:] = s0
    L = oracle_mod.lib()
    L.or_bssn_step_init(s.h)
    L.or_dust_step_init(s.h)
    L.or_dust_rk_evolve(s.h)
    v1 = np.exp(-10.0 * phi) * s0
    W = np.sqrt(1.0 + np.exp(-4.0 * phi) * s0 * s0 / (D0 * D0))
    dphi = -p * sigma4(k, h) * np.sin(k * x)
    force = 0.5 * W * D0 * np.exp(4.0 * phi) * v1 * v1 * 4.0 * dphi
    flux = s0 * v1
    dflux = (-np.roll(flux, -2, 0) + 8 * np.roll(flux, -1, 0) - 8 * np.roll(flux, 1, 0) + np.roll(flux, 2, 0)) / (12 * h)
    np.testing.assert_allclose(s.field("dust_S1_c"), dflux + force, rtol=1e-12, atol=1e-15)
    # D flux: alpha D v^1 with the same v^1
    fD = D0 * v1
    dfD = (-np.roll(fD, -2, 0) + 8 * np.roll(fD, -1, 0) - 8 * np.roll(fD, 1, 0) + np.roll(fD, 2, 0)) / (12 * h)
    np.testing.assert_allclose(s.field("dust_D_c"), dfD, rtol=1e-12, atol=1e-15)


def rk4_lagged(y0, f, h, nsteps):
    """The DustSim stage sequence on uniform data: stage n sees the source
    built from the state of stage n-1's input (stage 1: its own)."""
    y = np.array(y0, dtype=float)
    out = []
    for _ in range(nsteps):
        k1 = f(y, y)
        y1 = y + h / 2 * k1
        k2 = f(y1, y)
        y2 = y + h / 2 * k2
        k3 = f(y2, y1)
        y3 = y + h * k3
        k4 = f(y3, y2)
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(y.copy())
    return out


@pytest.mark.parametrize("lam", [0.0, 0.05])
def test_homogeneous_dust_lagged_rk4(oracle_mod, lam):
    n = 8
    s = dust_sim(oracle_mod, n, lam=lam)
    phi0, K0, D0 = 0.0, -3.0, 3.0 / (8 * math.pi)
    s.field("DIFFphi_p")[:] = phi0
    s.field("DIFFK_p")[:] = K0
    s.field("dust_D_p")[:] = D0
    h = s.p.dt

    def f(y, ylag):
        phi, K = y
        rho = D0 * math.exp(-6.0 * ylag[0]) + lam        # W = 1, detg from the lagged state
        S = -3.0 * lam
        return np.array([-K / 6.0, K * K / 3.0 + 4 * math.pi * (rho + S)])

    want = rk4_lagged([phi0, K0], f, h, 10)
    for step in range(10):
        s.run_steps(1)
        got = [s.field("DIFFphi_p")[2, 3, 4], s.field("DIFFK_p")[5, 1, 0]]
        np.testing.assert_allclose(got, want[step], rtol=1e-13)
        assert np.ptp(s.field("dust_D_p")) == 0 and s.field("dust_D_p")[0, 0, 0] == D0
    # the lag is visible: the unlagged sequence differs at O(dt^3) per step
    unlagged = rk4_lagged([phi0, K0], lambda y, yl: f(y, y), h, 10)
    assert abs(unlagged[-1][1] - want[-1][1]) > 1e3 * 1e-13 * abs(want[-1][1])


def test_total_D_conserved(oracle_mod):
    n = 12
    s = dust_sim(oracle_mod, n)
    rng = np.random.default_rng(5)
    shape = (n, n, n)
    x = np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")

    def smooth(a):
        return a * (np.cos(2 * np.pi * (X + 2 * Y) + rng.uniform(0, 6)) + np.sin(2 * np.pi * (Z - Y) + rng.uniform(0, 6)))

    for f in FIELDS[:6]:
        s.field(f + "_p")[:] = smooth(1e-3)
    s.field("DIFFphi_p")[:] = smooth(1e-3)
    s.field("DIFFK_p")[:] = -0.5 + smooth(1e-3)
    s.field("dust_D_p")[:] = 0.1 + smooth(1e-2)
    for a in range(3):
        s.field(f"dust_S{a + 1}_p")[:] = smooth(1e-3)
    tot0 = s.field("dust_D_p").sum()
    s.run_steps(3)
    tot = s.field("dust_D_p").sum()
    assert abs(tot - tot0) <= 1e-13 * abs(tot0)
    assert np.abs(s.field("dust_D_p") - 0.1).max() > 1e-3  # it did evolve
