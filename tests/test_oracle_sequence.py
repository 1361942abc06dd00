"""Pins the oracle's step sequencing (sims/static.cc:74-170, bssn.cc:219-305,
utils/RK4Register.h, utils/FRW.h): with spatially uniform data every stencil
vanishes, and the StaticSim sequence must equal the classical RK4 tableau
applied to the coupled ODE system (FRW background + homogeneous
perturbation) -- which holds only if each stage sees the FRW values of that
stage, the sources are rebuilt after every finalize, and the registers are
combined with the 1/6, 1/3, 1/3, 1/6 weights.  A source or FRW value taken
one stage late changes the result at O(dt^2)."""
import math

import numpy as np
import pytest


def rk4_reference(y0, f, h, nsteps):
    y = np.array(y0, dtype=float)
    out = []
    for _ in range(nsteps):
        k1 = f(y)
        k2 = f(y + h / 2 * k1)
        k3 = f(y + h / 2 * k2)
        k4 = f(y + h * k3)
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(y.copy())
    return out


@pytest.mark.parametrize("lam", [0.0, 0.02])
def test_homogeneous_static_dust_matches_rk4(oracle_mod, lam):
    n = 8
    rho0 = 3.0 / (8.0 * math.pi)
    s = oracle_mod.Sim(oracle_mod.make_params(n, order=4, lam=lam))
    dphi0, dK0, D = 2e-3, -5e-3, 4e-3
    s.field("DIFFphi_p")[:] = dphi0
    s.field("DIFFK_p")[:] = dK0
    s.field("DIFFD_a")[:] = D
    s.frw_set(0.0, -3.0)
    s.frw_add_fluid(rho0, 0.0)
    h = s.p.dt

    def f(y):
        phiF, KF, rhoF, dphi, dK = y
        DIFFr = math.exp(-6 * (dphi + phiF)) * D + rhoF * math.expm1(-6 * dphi) + lam
        DIFFS = -3 * lam
        return np.array([-KF / 6, KF * KF / 3 + 4 * math.pi * rhoF, KF * rhoF,
                         -dK / 6, dK * (dK + 2 * KF) / 3 + 4 * math.pi * (DIFFr + DIFFS)])

    want = rk4_reference([0.0, -3.0, rho0, dphi0, dK0], f, h, 12)
    for step in range(12):
        s.run_steps(1)
        phiF, KF, rhoF, _ = s.frw_get()
        got = [phiF, KF, rhoF, s.field("DIFFphi_p")[1, 2, 3], s.field("DIFFK_p")[4, 0, 7]]
        np.testing.assert_allclose(got, want[step], rtol=1e-13, atol=1e-17)
    # uniform data stays uniform; the other fields stay zero up to the
    # roundoff of the stencil sums on uniform data (the reference's stencils
    # do not cancel exactly either, utils/math.h:231-263)
    assert np.ptp(s.field("DIFFK_p")) == 0
    for f_ in ("A11_p", "DIFFgamma12_p", "Gamma3_p", "DIFFalpha_p"):
        assert np.abs(s.field(f_)).max() < 1e-15


def test_homogeneous_conformal_flrw_lapse_matches_rk4(oracle_mod):
    """ConformalFLRW lapse (BSSNGaugeHandler.cc:24-27): d_t DIFFalpha =
    -alpha K_avg / 3, with K_avg the conformal average of DIFFK over the
    stage's _a (setExtraFieldData after the previous finalize, bssn.cc:165-170).
    On uniform data K_avg equals DIFFK of the stage input, so the sequence is
    classical RK4 of the 6-variable ODE; a K_avg one stage stale, or the full
    K averaged instead of the perturbation, fails at O(dt^2) or O(1)."""
    n = 8
    rho0 = 3.0 / (8.0 * math.pi)
    s = oracle_mod.Sim(oracle_mod.make_params(n, order=4, lapse="ConformalFLRW"))
    dphi0, dK0, da0, D = 2e-3, -5e-3, 1e-3, 4e-3
    s.field("DIFFphi_p")[:] = dphi0
    s.field("DIFFK_p")[:] = dK0
    s.field("DIFFalpha_p")[:] = da0
    s.field("DIFFD_a")[:] = D
    s.frw_set(0.0, -3.0)
    s.frw_add_fluid(rho0, 0.0)
    h = s.p.dt

    def f(y):
        phiF, KF, rhoF, dphi, dK, da = y
        alpha = 1.0 + da
        DIFFr = math.exp(-6 * (dphi + phiF)) * D + rhoF * math.expm1(-6 * dphi)
        return np.array([-KF / 6, KF * KF / 3 + 4 * math.pi * rhoF, KF * rhoF,
                         -(alpha * dK + da * KF) / 6,
                         alpha * dK * (dK + 2 * KF) / 3 + 4 * math.pi * alpha * DIFFr + 4 * math.pi * da * rhoF,
                         -alpha * dK / 3])

    want = rk4_reference([0.0, -3.0, rho0, dphi0, dK0, da0], f, h, 8)
    for step in range(8):
        s.run_steps(1)
        phiF, KF, rhoF, _ = s.frw_get()
        got = [phiF, KF, rhoF, s.field("DIFFphi_p")[1, 2, 3], s.field("DIFFK_p")[4, 0, 7],
               s.field("DIFFalpha_p")[5, 5, 1]]
        np.testing.assert_allclose(got, want[step], rtol=1e-13, atol=1e-17)
        # setExtraFieldData after K4Finalize: averages of the new state
        ex = s.extra()
        dK, dphi = want[step][4], want[step][3]
        assert ex["K_min"] == s.field("DIFFK_p").min()
        assert abs(ex["K_avg"] - dK) <= 1e-13 * abs(dK)
        assert abs(ex["avg_vol"] - math.exp(6 * dphi)) <= 1e-13


# d_t DIFFalpha of the shift-free lapse gauges (BSSNGaugeHandler.cc), as
# functions of alpha, the full K, DIFFK, K_avg (= DIFFK on uniform data),
# K_FRW and the full phi
LAPSES = {
    "Harmonic": lambda a, K, dK, Kav, KF, ph: -a * K,
    "Anharmonic": lambda a, K, dK, Kav, KF, ph: a * a * (K - Kav),
    "OnePlusLog": lambda a, K, dK, Kav, KF, ph: -2.0 * a * (K - Kav) * 0.7,        # gd_c = 0.7
    "ConformalSync": lambda a, K, dK, Kav, KF, ph: -a * KF / 3.0,
    "AwAGaugeWave": lambda a, K, dK, Kav, KF, ph: -a * a * dK,
    "AwAShiftedWave": lambda a, K, dK, Kav, KF, ph: -dK,
    "DampedWave": lambda a, K, dK, Kav, KF, ph: a * a * (0.3 * (12.0 * ph * 0.5 - math.log(a)) - K),
    "TestAijDriverLapse": lambda a, K, dK, Kav, KF, ph: 0.0,
}


@pytest.mark.parametrize("lapse", sorted(LAPSES))
def test_homogeneous_lapse_gauges_match_rk4(oracle_mod, lapse):
    """Every lapse gauge on uniform data: classical RK4 of the background +
    perturbation ODE with d_t DIFFalpha from the gauge.  Pins which K each
    gauge reads (full K, DIFFK, the averaged perturbation, K_FRW), the
    gauge parameters and that every stage sees its own lapse and K_avg."""
    n = 8
    rho0 = 3.0 / (8.0 * math.pi)
    s = oracle_mod.Sim(oracle_mod.make_params(n, order=4, lapse=lapse, gd_c=0.7, dw_mu_l=0.3, dw_p=0.5))
    dphi0, dK0, da0, D = 2e-3, -5e-3, 1e-3, 4e-3
    s.field("DIFFphi_p")[:] = dphi0
    s.field("DIFFK_p")[:] = dK0
    s.field("DIFFalpha_p")[:] = da0
    s.field("DIFFD_a")[:] = D
    s.frw_set(0.0, -3.0)
    s.frw_add_fluid(rho0, 0.0)
    h = s.p.dt
    fl = LAPSES[lapse]

    def f(y):
        phiF, KF, rhoF, dphi, dK, da = y
        alpha = 1.0 + da
        DIFFr = math.exp(-6 * (dphi + phiF)) * D + rhoF * math.expm1(-6 * dphi)
        return np.array([-KF / 6, KF * KF / 3 + 4 * math.pi * rhoF, KF * rhoF,
                         -(alpha * dK + da * KF) / 6,
                         alpha * dK * (dK + 2 * KF) / 3 + 4 * math.pi * alpha * DIFFr + 4 * math.pi * da * rhoF,
                         fl(alpha, dK + KF, dK, dK, KF, dphi + phiF)])

    want = rk4_reference([0.0, -3.0, rho0, dphi0, dK0, da0], f, h, 6)
    for step in range(6):
        s.run_steps(1)
        phiF, KF, rhoF, _ = s.frw_get()
        got = [phiF, KF, rhoF, s.field("DIFFphi_p")[1, 2, 3], s.field("DIFFK_p")[4, 0, 7],
               s.field("DIFFalpha_p")[5, 5, 1]]
        np.testing.assert_allclose(got, want[step], rtol=1e-12, atol=1e-17)
