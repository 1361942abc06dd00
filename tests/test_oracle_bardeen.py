"""Pins the Bardeen / SVT oracle (oracle/bardeen.py, a restatement of
components/bssn/bardeen.cc) without the reference, which needs FFTW:

  * its stencils are the C oracle's bit for bit (those are bitwise equal to
    utils/math.h built from source, tests/test_oracle_ref.py);
  * the inverse Laplacian of a single Fourier mode is the mode over
    -|2 pi k / L|^2, and the mean is removed;
  * closed forms of the SVT decomposition: for a metric perturbation
    h_ij = A delta_ij + d_i d_j B (scalar), d_i C_j + d_j C_i with
    d_i C_i = 0 (vector) and a transverse-traceless D_ij (tensor), on a
    static background (a = 1, H = 0), the recovered potentials converge to
    the ones put in at the order of the stencils (exactly for the tensor
    mode, whose derivatives the decomposition never takes); the Bardeen
    potentials of the scalar mode are Psi = -A/2, Phi = -B''/2, and the
    time derivatives of A and B are recovered from those of h;
  * d^2/dt^2 of the conformal metric and of phi (bardeen.cc:102-143) equal
    the time derivative of the C oracle's right-hand side along an RK4
    trajectory (central difference), with a live lapse;
  * the linearised-constraint violation (bardeen.cc:412-415) is the
    no-anisotropic-stress condition 2 (Phi - Psi) of the Bardeen potentials
    formed separately (bardeen.cc:329-340), at every point.
"""
import math

import numpy as np
import pytest

import bardeen as bd

FIELDS = ["DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23", "DIFFgamma33", "DIFFphi",
          "A11", "A12", "A13", "A22", "A23", "A33", "DIFFK", "Gamma1", "Gamma2", "Gamma3", "DIFFalpha"]
SYM = bd.SYM


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_stencils_equal_c_oracle(oracle_mod, order):
    shape = (9, 10, 11)
    f = np.ascontiguousarray(np.random.default_rng(order).standard_normal(shape))
    dx = 1.0 / 9
    s = oracle_mod.Sim(oracle_mod.make_params(shape, order=order, dx=dx))
    L = oracle_mod.lib()
    fp = oracle_mod.dptr(f)
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                for d in range(3):
                    assert bd.d1(f, d, order, dx)[i, j, k] == L.or_stencil(s.h, fp, 0, d + 1, 0, i, j, k)
                for p, q in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2), (1, 0), (2, 1)]:
                    assert bd.dd(f, p, q, order, dx)[i, j, k] == L.or_stencil(s.h, fp, 1, p + 1, q + 1, i, j, k)
        break  # one i-plane is enough for every stencil tap (all planes wrap the same way)


def test_inverse_laplacian_single_mode():
    n, box = 16, 2.0
    x = np.arange(n) * box / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    k = np.array([2, -1, 3])
    f = np.cos(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) / box + 0.3) + 0.7
    got = bd.inverse_laplacian(f, box)
    want = -(f - 0.7) / (np.sum(k * k) * (2 * np.pi / box) ** 2)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


def static_background(shape):
    z = np.zeros(shape)
    a_reg = {f: z.copy() for f in FIELDS}
    c_reg = {f: z.copy() for f in FIELDS}
    return a_reg, c_reg


def grid(n):
    x = np.arange(n) / n
    return np.meshgrid(x, x, x, indexing="ij")


def scalar_mode(n, order):
    """h_ij = A0 delta_ij + d_i d_j B0 (continuous derivatives), its first
    and second time derivatives through dA, dB and ddA, ddB factors."""
    X, Y, Z = grid(n)
    th = 2 * np.pi * (1 * X + 2 * Y - 1 * Z) + 0.4
    k = 2 * np.pi * np.array([1, 2, -1])
    A0 = 1e-3 * np.sin(2 * np.pi * (X - Z) + 0.1)
    B0 = 2e-4 * np.cos(th)
    dA, dB = 0.7, -1.3         # d_t A0 = dA A0, d_t B0 = dB B0
    ddA, ddB = 0.2, 0.5
    a_reg, c_reg = static_background((n, n, n))
    for I, J in SYM:
        dd_B = -k[I] * k[J] * B0
        s = (A0 if I == J else 0.0) + dd_B
        a_reg[f"DIFFgamma{I + 1}{J + 1}"] = s
        c_reg[f"DIFFgamma{I + 1}{J + 1}"] = dA * (A0 if I == J else 0.0) + dB * dd_B
        # d2t gamma = -2 alpha ev_A at alpha = 1, d_t alpha = 0
        c_reg[f"A{I + 1}{J + 1}"] = -0.5 * (ddA * (A0 if I == J else 0.0) + ddB * dd_B)
    return a_reg, c_reg, A0, B0, (dA, dB, ddA, ddB)


@pytest.mark.parametrize("order,ratio,bound", [(4, 12.0, 5e-3), (6, 40.0, 1e-4)])
def test_scalar_mode_converges(order, ratio, bound):
    """Errors relative to the amplitudes put in (measured: 1.2e-3 at order 4
    and 1.9e-5 at order 6 on 64^3, ratios 15.7 and 62 from 32^3; the d_i d_j B
    part of the metric is 50 times the A part, so A carries most of it)."""
    errs = []
    for n in (32, 64):
        a_reg, c_reg, A0, B0, (dA, dB, ddA, ddB) = scalar_mode(n, order)
        f, v, _ = bd.set_potentials(a_reg, c_reg, 0.0, order, 1.0 / n, 1.0, use_ml=False)
        e = max(np.abs(f["A"] - A0).max() / np.abs(A0).max(), np.abs(f["B"] - B0).max() / np.abs(B0).max(),
                np.abs(f["dt_A"] - dA * A0).max() / np.abs(A0).max(),
                np.abs(f["dt_B"] - dB * B0).max() / np.abs(B0).max(),
                np.abs(f["d2t_A"] - ddA * A0).max() / np.abs(A0).max(),
                np.abs(f["d2t_B"] - ddB * B0).max() / np.abs(B0).max(),
                np.abs(f["Psi"] + 0.5 * A0).max() / np.abs(A0).max(),
                np.abs(f["Phi"] + 0.5 * ddB * B0).max() / np.abs(B0).max(),
                max(np.abs(f[c]).max() for c in ("C1", "C2", "C3", "D11", "D12", "D23")) / np.abs(A0).max())
        errs.append(e)
    assert errs[1] < bound and errs[0] / errs[1] > ratio, errs


def test_vector_mode_converges():
    """Errors relative to the amplitude of C (measured 1.8e-2 at 32^3 and
    1.1e-3 at 64^3: the mode has 4 points per wavelength along y at 16^3)."""
    errs = []
    for n in (32, 64):
        X, Y, Z = grid(n)
        Cx = 1e-3 * np.sin(2 * np.pi * (Z + 2 * Y) + 0.2)     # d_i C_i = 0: C_x depends on y, z only
        dyC, dzC = 1e-3 * 4 * np.pi * np.cos(2 * np.pi * (Z + 2 * Y) + 0.2), 1e-3 * 2 * np.pi * np.cos(
            2 * np.pi * (Z + 2 * Y) + 0.2)
        a_reg, c_reg = static_background((n, n, n))
        a_reg["DIFFgamma12"] = dyC
        a_reg["DIFFgamma13"] = dzC
        f, _, _ = bd.set_potentials(a_reg, c_reg, 0.0, 4, 1.0 / n, 1.0, use_ml=False)
        errs.append(max(np.abs(f["C1"] - Cx).max(), np.abs(f["C2"]).max(), np.abs(f["C3"]).max(),
                        np.abs(f["A"]).max(), np.abs(f["B"]).max(), np.abs(f["D12"]).max()) / 1e-3)
    assert errs[1] < 3e-3 and errs[0] / errs[1] > 12.0, errs


def test_tensor_mode_exact():
    n = 16
    X, Y, Z = grid(n)
    hp, hx = 1e-3 * np.cos(2 * np.pi * 2 * Z + 0.3), 5e-4 * np.sin(2 * np.pi * Z)
    a_reg, c_reg = static_background((n, n, n))
    a_reg["DIFFgamma11"], a_reg["DIFFgamma22"], a_reg["DIFFgamma12"] = hp, -hp, hx
    c_reg["DIFFgamma11"], c_reg["DIFFgamma22"], c_reg["DIFFgamma12"] = 0.3 * hp, -0.3 * hp, 0.3 * hx
    f, v, _ = bd.set_potentials(a_reg, c_reg, 0.0, 4, 1.0 / n, 1.0, use_ml=False)
    # zero up to the rounding of h_ij = (1 + h) - 1 (amplitudes 1e-3)
    for name in ("A", "B", "dt_A", "dt_B", "C1", "C2", "C3", "Phi", "Psi", "F", "E"):
        assert np.abs(f[name]).max() < 1e-15, name
    np.testing.assert_allclose(f["D11"], hp, rtol=0, atol=1e-15)
    np.testing.assert_allclose(f["D22"], -hp, rtol=0, atol=1e-15)
    np.testing.assert_allclose(f["D12"], hx, rtol=0, atol=1e-15)
    assert np.abs(f["D13"]).max() < 1e-15 and np.abs(f["D33"]).max() < 1e-15


def test_matter_lambda_scale_factor():
    """bardeen.h:176-229: matter + Lambda FLRW with a(0) = 1, H(0) = 1:
    a^3 = cosh(3 sqrt(OL) t / 2) + sinh(...) / sqrt(OL) solves
    (a'/a)^2 = (1 - OL) a^-3 + OL; Omega_L = 0 is a = (1 + 3t/2)^(2/3)."""
    for OL in (0.0, 0.3, 0.7):
        for t in (0.0, 0.4, 1.3):
            a = bd.ml_scale_factor(t, OL)
            e = 1e-6
            dadt = (bd.ml_scale_factor(t + e, OL) - bd.ml_scale_factor(t - e, OL)) / (2 * e)
            d2 = (bd.ml_scale_factor(t + e, OL) - 2 * a + bd.ml_scale_factor(t - e, OL)) / e / e
            assert abs(dadt / a - bd.ml_hubble(t, OL)) < 1e-8
            assert abs(d2 - bd.ml_d2adt2(t, OL)) < 2e-3 * max(1, abs(d2))
    assert abs(bd.ml_scale_factor(0.5, 0.0) - (1 + 0.75) ** (2 / 3)) < 1e-15


def test_second_time_derivatives_follow_the_trajectory(oracle_mod):
    """d2t gamma~_ij and d2t phi (bardeen.cc:102-143, no shift, live lapse)
    equal d/dt of the oracle's right-hand side, by a central difference of
    RK4 sub-steps of +-h around the state (error O(h^2))."""
    po = oracle_mod
    n = 12
    rng = np.random.default_rng(3)
    X, Y, Z = grid(n)
    state = {}
    for q, f in enumerate(FIELDS):
        state[f] = 2e-3 * np.sin(2 * np.pi * (X + (q % 3) * Y - Z) + q) * (0.5 + 0.5 * rng.random())
    state["DIFFalpha"] = 0.02 * np.cos(2 * np.pi * (Y + Z))

    def rhs_at(h_steps, h):
        s = po.Sim(po.make_params(n, order=4, lapse="Harmonic", matter="none", dt_frac=h * n))
        for f in FIELDS:
            s.field(f + "_p")[:] = state[f]
        for _ in range(h_steps if h_steps > 0 else 0):
            s.run_steps(1)
        s.stepInit()
        s.clearSrc()
        s.RKEvolve()
        return {f: s.field(f + "_a").copy() for f in FIELDS}, {f: s.field(f + "_c").copy() for f in FIELDS}

    h = 1e-4
    base = po.Sim(po.make_params(n, order=4, lapse="Harmonic", matter="none"))
    for f in FIELDS:
        base.field(f + "_p")[:] = state[f]
    base.stepInit()  # start on the normalised state (det gamma~ = 1, A_ij trace-free)
    for f in FIELDS:
        state[f] = base.field(f + "_a").copy()
    a0, c0 = rhs_at(0, h)
    s = po.Sim(po.make_params(n, order=4, lapse="Harmonic", matter="none", dt_frac=-h * n))
    for f in FIELDS:
        s.field(f + "_p")[:] = state[f]
    s.run_steps(1)
    s.stepInit(); s.clearSrc(); s.RKEvolve()
    cm = {f: s.field(f + "_c").copy() for f in FIELDS}
    _, cp = rhs_at(1, h)
    _, _, inter = bd.set_potentials(a0, c0, 0.0, 4, 1.0 / n, 1.0, use_ml=False)
    for I, J in SYM:
        name = f"DIFFgamma{I + 1}{J + 1}"
        want = (cp[name] - cm[name]) / (2 * h)
        np.testing.assert_allclose(inter[f"d2t_g{I + 1}{J + 1}"], want, rtol=0, atol=1e-6 * np.abs(want).max())
    want = (cp["DIFFphi"] - cm["DIFFphi"]) / (2 * h)
    np.testing.assert_allclose(inter["d2t_phi"], want, rtol=0, atol=1e-6 * np.abs(want).max())
    # h_ij, dt h_ij: e^{4 phi} gamma~ minus the background, and its time derivative
    for I, J in SYM:
        g = a0[f"DIFFgamma{I + 1}{J + 1}"] + (1.0 if I == J else 0.0)
        gp = {k: v for k, v in cp.items()}
        np.testing.assert_allclose(inter[f"h{I + 1}{J + 1}"] + (inter["a"] ** 2 if I == J else 0.0),
                                   np.exp(4 * a0["DIFFphi"]) * g, rtol=1e-15, atol=0)


@pytest.mark.parametrize("use_ml", [True, False])
def test_linear_violation_is_phi_minus_psi(oracle_mod, use_ml):
    """lin_viol = E + A - a^2 B'' - 3 a a' B' + 2 a F' + 4 a' F equals
    2 (Phi - Psi) with Phi = (E + 2(a' F + a F') - 2 a a' B' - a^2 B'')/2 and
    Psi = -(A + 2 H (a F - a^2 B'/2))/2: a dropped term or a wrong factor in
    either formula breaks the identity.  Data with a shift, so that F != 0."""
    po = oracle_mod
    n = 12
    X, Y, Z = grid(n)
    a_reg, c_reg = static_background((n, n, n))
    for q, f in enumerate(FIELDS):
        a_reg[f] = 1e-3 * np.sin(2 * np.pi * (X + q * Y) + q) + (0.02 if f == "DIFFK" else 0.0)
        c_reg[f] = 2e-3 * np.cos(2 * np.pi * (Z - q * X) + 0.3 * q)
    for m in range(3):
        a_reg[f"beta{m + 1}"] = 1e-2 * np.sin(2 * np.pi * (Y + m * Z) + m)
        c_reg[f"beta{m + 1}"] = 3e-3 * np.cos(2 * np.pi * (X - m * Y))
    f, v, inter = bd.set_potentials(a_reg, c_reg, 0.3, 4, 1.0 / n, 1.0, use_ml=use_ml, omega_L=0.4)
    lv = inter["lin_viol"]
    assert np.abs(f["F"]).max() > 1e-6 and np.abs(lv).max() > 1e-5
    np.testing.assert_allclose(lv, 2.0 * (f["Phi"] - f["Psi"]), rtol=0, atol=1e-13 * np.abs(lv).max())
    # the signed maxima among the 17 violations (bardeen.cc:422-444)
    assert v[7] == float(np.max(lv))
    assert v[3] == float(np.max(inter["lin_viol_der"]))
