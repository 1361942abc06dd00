"""Pins the oracle's shift build (USE_BSSN_SHIFT, cosmo_macros.h:43-44) at
the grid level, where the point algebra test (test_oracle_adm.py) cannot
reach: the upwind advection stencils wired into every field, the shift
gauges of BSSNGaugeHandler.cc:113-273 and the TestKDriverLapse (:93-102).

Exact solutions.  The AwA gauge waves (components/bssn/bssn_ic.cc:179-360)
are flat spacetime in coordinates where every variable is a function of
x - t alone, so each field's exact time derivative is minus its x
derivative:
  * the shifted gauge wave, H = A sin 2 pi (x - t): gamma_xx = 1 + H,
    alpha = (1 + H)^(-1/2), beta^x = -H / (1 + H).  It is a solution of the
    harmonic lapse alpha^2 K (the reference's AwAGaugeWave lapse) and of the
    reference's AwAShiftedWave shift -2 K alpha;
  * the gauge wave, H = A sin 2 pi (x - t): gamma_xx = 1 - H,
    alpha = (1 - H)^(1/2), beta = 0.  It is a solution of the damped-wave
    gauge with mu_L = mu_S = 0 (harmonic lapse and shift), along x and z;
    along y the reference's DampedWaveShift2 reads Gamma3 for Gamma2
    (BSSNGaugeHandler.cc:185), so the y wave is not kept.
The grid right-hand side must converge to those exact derivatives at the
order of its stencils: 4 with the 4th-order centred and lopsided stencils
(math.h:231-337), 1 at order 2, whose upwind stencil adds -dx/2 times a
second difference (math.h:1087-1130).

Closed forms on homogeneous data:
  * TestKDriverLapse: with K_avg = DIFFK and rho_avg = DIFFr the driver's
    terms cancel the physical ones, d_t alpha = -k K_FRW (1 + 2 DIFFK/3);
  * AijDriverShift: d_t A_ij = alpha (K A_ij - 2 A_ik A^k_j) with no
    gradients, so d_t beta^i = -k sum_j |K A_ij - 2 (A A)_ij + A_ij|.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
from torch.func import jacfwd, vmap  # noqa: E402

torch.set_default_dtype(torch.float64)
PI = math.pi
SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def wave_fields(x, amp, shifted, d):
    """BSSN variables of the (shifted) gauge wave along axis d (0..2) at
    position x along it and t = 0 (bssn_ic.cc:193-259, 268-360)."""
    H = amp * torch.sin(2 * PI * x)
    dtH = -2 * PI * amp * torch.cos(2 * PI * x)
    s = 1.0 if shifted else -1.0
    q = 1 + s * H
    Kxx = dtH / (2 * torch.sqrt(q))
    K = Kxx / q
    f = {"DIFFphi": torch.log(q) / 12, "DIFFK": K,
         "DIFFalpha": (q ** -0.5 if shifted else q ** 0.5) - 1.0}
    for a, b in SYM:
        on = a == b == d
        f[f"DIFFgamma{a + 1}{b + 1}"] = (q ** (2 / 3) - 1 if on else q ** (-1 / 3) - 1) if a == b else 0 * x
        f[f"A{a + 1}{b + 1}"] = ((Kxx * q ** (-1 / 3) - K / 3 * q ** (2 / 3)) if on else -K / 3 * q ** (-1 / 3)) \
            if a == b else 0 * x
    for i in range(3):
        f[f"Gamma{i + 1}"] = (-s * 2 / 3 * dtH * q ** (-5 / 3)) if i == d else 0 * x
        f[f"beta{i + 1}"] = (-H / (1 + H)) if (shifted and i == d) else 0 * x
    return f


def load_wave(po, n, order, amp, shifted, d, lapse, shift, **kw):
    shape = [4, 4, 4]
    shape[d] = n
    s = po.Sim(po.make_params(tuple(shape), order=order, dx=1.0 / n, lapse=lapse, shift=shift, matter="none",
                              gauge_wave_dir=d + 1, **kw))
    x = torch.arange(n, dtype=torch.float64) / n
    vals = wave_fields(x, amp, shifted, d)
    view = [1, 1, 1]
    view[d] = n
    for name, v in vals.items():
        s.field(name + "_p")[:] = np.broadcast_to(v.numpy().reshape(view), shape)
    return s, shape


def exact_dt(n, amp, shifted, d, names):
    """-d/dx of each field (every field is a function of x - t)."""
    x = torch.arange(n, dtype=torch.float64) / n
    out = {}
    for name in names:
        g = vmap(jacfwd(lambda y: wave_fields(y, amp, shifted, d)[name]))(x)
        out[name] = -g.numpy()
    return out


def rhs_error(po, n, order, amp, shifted, d, lapse, shift, names, **kw):
    s, shape = load_wave(po, n, order, amp, shifted, d, lapse, shift, **kw)
    s.stepInit()
    s.clearSrc()
    s.RKEvolve()
    want = exact_dt(n, amp, shifted, d, names)
    idx = [0, 0, 0]
    err = 0.0
    for name in names:
        got = s.field(name + "_c")
        line = [got[tuple(idx[:d] + [k] + idx[d + 1:])] for k in range(n)]
        err = max(err, float(np.abs(np.array(line) - want[name]).max()))
    return err, s


BSSN = ["DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23", "DIFFgamma33", "DIFFphi",
        "A11", "A12", "A13", "A22", "A23", "A33", "DIFFK", "Gamma1", "Gamma2", "Gamma3", "DIFFalpha"]


@pytest.mark.parametrize("d", [0, 1, 2])
def test_shifted_gauge_wave_converges(oracle_mod, d):
    """Order 4: every field including beta^d converges at 4th order to the
    exact x - t derivatives (upwind advection of all 18 fields, d beta
    couplings, second derivatives of beta in d_t Gt^i, the shift's own
    equation)."""
    names = BSSN + [f"beta{i + 1}" for i in range(3)]
    e1, _ = rhs_error(oracle_mod, 32, 4, 0.1, True, d, "AwAGaugeWave", "AwAShiftedWave", names)
    e2, _ = rhs_error(oracle_mod, 64, 4, 0.1, True, d, "AwAGaugeWave", "AwAShiftedWave", names)
    assert e2 < 2e-4 and e1 / e2 > 12.0, (e1, e2)


def test_shifted_gauge_wave_order2_is_first_order(oracle_mod):
    """Order 2: the upwind stencil carries -dx/2 times a one-sided second
    difference (math.h:1087-1130), so advection is first-order accurate."""
    names = BSSN + ["beta1"]
    e1, _ = rhs_error(oracle_mod, 64, 2, 0.1, True, 0, "AwAGaugeWave", "AwAShiftedWave", names)
    e2, _ = rhs_error(oracle_mod, 128, 2, 0.1, True, 0, "AwAGaugeWave", "AwAShiftedWave", names)
    assert 1.6 < e1 / e2 < 2.6, (e1, e2)


@pytest.mark.parametrize("d", [0, 2])
def test_gauge_wave_damped_wave_gauge(oracle_mod, d):
    """The damped-wave gauge with mu_L = mu_S = 0 keeps the gauge wave: the
    lapse -alpha^2 K and a shift whose right-hand side vanishes (beta^2
    excepted along z, where the shift's Gamma2 is read as Gamma3)."""
    names = BSSN + [f"beta{i + 1}" for i in range(3) if not (d == 2 and i == 1)]
    e1, s = rhs_error(oracle_mod, 32, 4, 0.1, False, d, "DampedWave", "DampedWave", names)
    e2, _ = rhs_error(oracle_mod, 64, 4, 0.1, False, d, "DampedWave", "DampedWave", names)
    assert e2 < 2e-4 and e1 / e2 > 12.0, (e1, e2)


@pytest.mark.parametrize("d", [1, 2])
def test_gauge_wave_damped_wave_shift_quirk(oracle_mod, d):
    """DampedWaveShift2 reads Gamma3 where Gamma2 is meant
    (BSSNGaugeHandler.cc:185).  Along y (Gt^3 = 0) d_t beta^2 converges to
    -alpha^2 e^{-4 phi} Gt^2 instead of 0; along z (Gt^2 = 0, no y
    gradients) to +alpha^2 e^{-4 phi} Gt^3."""
    n = 64
    _, s = rhs_error(oracle_mod, n, 4, 0.1, False, d, "DampedWave", "DampedWave", ["beta2"])
    x = torch.arange(n, dtype=torch.float64) / n
    f = wave_fields(x, 0.1, False, d)
    al = (f["DIFFalpha"] + 1).numpy()
    g = f["Gamma2"].numpy() if d == 1 else -f["Gamma3"].numpy()
    want = -al * al * np.exp(-4 * f["DIFFphi"].numpy()) * g
    got = s.field("beta2_c")[0, :, 0] if d == 1 else s.field("beta2_c")[0, 0, :]
    assert np.abs(want).max() > 0.1
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-3 * np.abs(want).max())


def test_gamma_driver(oracle_mod):
    """Gamma driver (USE_GAMMA_DRIVER): d_t beta^i = B^i, d_t B^i =
    3/4 d_t Gt^i - eta B^i.  On the gauge wave with B = 0 the shift stays
    put and d_t B^x converges to 3/4 of the exact d_t Gt^x."""
    names = ["Gamma1", "beta1", "beta2", "beta3"]
    errs = []
    for n in (32, 64):
        err, s = rhs_error(oracle_mod, n, 4, 0.1, False, 0, "AwAGaugeWave", "GammaDriver", names, gd_eta=0.7)
        want = 0.75 * exact_dt(n, 0.1, False, 0, ["Gamma1"])["Gamma1"]
        errs.append(max(err, float(np.abs(s.field("auxB1_c")[:, 0, 0] - want).max())))
        assert np.abs(s.field("auxB2_c")).max() < 1e-12 and np.abs(s.field("auxB3_c")).max() < 1e-12
    assert errs[1] < 2e-4 and errs[0] / errs[1] > 12.0, errs
    # with B != 0: d_t beta = B, and the damping -eta B
    s, _ = load_wave(oracle_mod, 16, 4, 0.1, False, 0, "AwAGaugeWave", "GammaDriver", gd_eta=0.7)
    s.field("auxB2_p")[:] = 0.3
    s.stepInit()
    s.clearSrc()
    s.RKEvolve()
    assert np.all(s.field("beta2_c") == 0.3)
    np.testing.assert_allclose(s.field("auxB2_c"), 0.75 * s.field("Gamma2_c") - 0.7 * 0.3, rtol=0, atol=1e-15)


def homogeneous(po, lapse, shift, dK, A, rho=0.0, k=0.04):
    s = po.Sim(po.make_params(8, order=4, lapse=lapse, shift=shift, matter="none", k_driver_coeff=k))
    s.frw_set(0.0, -3.0)
    s.frw_add_fluid(3 / (8 * PI), 0.0)
    s.field("DIFFK_p")[:] = dK
    for q, (a, b) in enumerate(SYM):
        s.field(f"A{a + 1}{b + 1}_p")[:] = A[a][b]
    s.field("DIFFr_a")[:] = rho
    s.stepInit()
    s.RKEvolve()
    return s


def test_k_driver_lapse_closed_form(oracle_mod):
    dK, k = 0.13, 0.04
    s = homogeneous(oracle_mod, "TestKDriverLapse", "none", dK, np.zeros((3, 3)), rho=0.05, k=k)
    ex = s.extra()
    assert ex["K_avg"] == pytest.approx(dK, abs=1e-15) and ex["rho_avg"] == pytest.approx(0.05, abs=1e-15)
    got = s.field("DIFFalpha_c")
    want = -k * (-3.0) * (1 + 2 * dK / 3)
    np.testing.assert_allclose(got, want, rtol=1e-13)


def test_aij_driver_shift_closed_form(oracle_mod):
    a, b, c, e = 0.11, -0.04, 0.03, -0.02
    A = np.array([[a, c, e], [c, b, 0.0], [e, 0.0, -a - b]])
    dK, k = 0.07, 0.05
    s = homogeneous(oracle_mod, "Static", "AijDriverShift", dK, A, k=k)
    Kf = dK - 3.0
    ev = Kf * A - 2 * A @ A
    for i in range(3):
        want = -k * np.abs(ev[i] + A[i]).sum()
        np.testing.assert_allclose(s.field(f"beta{i + 1}_c"), want, rtol=1e-12)
