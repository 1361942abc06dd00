"""The oracle against the reference's own code, value by value.

Pieces of the reference that build from their own sources (oracle/_ref:
utils/math.h, utils/RK4Register.h, utils/FRW.h, components/static/static.cc,
tests/rk4.cc, tests/reference_frw.cc) were run to produce
tests/golden/ref_vectors.json (tests/golden/make_golden.py).  The oracle must
reproduce every vector bitwise.  Where oracle/_ref is present the same
comparison also runs live, and the reference's known-answer tests run.
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def field_input(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def oracle_stencils(po, order, f, ko):
    nx, ny, nz = f.shape
    s = po.Sim(po.make_params((nx, ny, nz), order=order, ko=ko))
    L = po.lib()
    fp = po.dptr(np.ascontiguousarray(f))
    d1 = np.zeros((3, nx, ny, nz))
    d2 = np.zeros((9, nx, ny, nz))
    kk = np.zeros((nx, ny, nz))
    P = [(1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3), (2, 1), (3, 1), (3, 2)]
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for d in range(3):
                    d1[d, i, j, k] = L.or_stencil(s.h, fp, 0, d + 1, 0, i, j, k)
                for q, (a, b) in enumerate(P):
                    d2[q, i, j, k] = L.or_stencil(s.h, fp, 1, a, b, i, j, k)
                kk[i, j, k] = L.or_stencil(s.h, fp, 2, 0, 0, i, j, k)
    return s, d1, d2, kk


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_stencils_bitwise_vs_reference(oracle_mod, order):
    shape = tuple(GOLD["grid"])
    f = field_input(shape, GOLD["seed_field"] + order)
    s, d1, d2, kk = oracle_stencils(oracle_mod, order, f, GOLD["ko_coeff"])
    ref = GOLD["stencils"][str(order)]
    assert digest(d1) == ref["derivative"]["sha256"]
    assert digest(d2) == ref["double_derivative"]["sha256"]
    assert digest(kk) == ref["ko"]["sha256"]
    # KO is implemented only at orders 2 and 8 (utils/math.h:36-72)
    assert (np.abs(kk).max() > 0) == (order in (2, 8))


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_reductions_bitwise_vs_reference(oracle_mod, order):
    shape = tuple(GOLD["grid"])
    f = np.ascontiguousarray(field_input(shape, GOLD["seed_field"] + order))
    g = np.ascontiguousarray(0.01 * field_input(shape, GOLD["seed_field"] + 100 + order))
    s = oracle_mod.Sim(oracle_mod.make_params(shape, order=order))
    L = oracle_mod.lib()
    got = [L.or_reduction(s.h, kind, oracle_mod.dptr(f), oracle_mod.dptr(g), 0.0123) for kind in range(7)]
    want = [float.fromhex(x) for x in GOLD["stencils"][str(order)]["reductions"]]
    assert got == want


def test_static_source_bitwise_vs_reference(oracle_mod):
    shape = tuple(GOLD["grid"])
    g = np.ascontiguousarray(0.01 * field_input(shape, GOLD["seed_field"] + 104))
    D = np.ascontiguousarray(0.05 * field_input(shape, GOLD["seed_field"] + 7))
    gs = GOLD["static_src"]
    s = oracle_mod.Sim(oracle_mod.make_params(shape, order=4))
    s.field("DIFFphi_a")[:] = g
    s.field("DIFFD_a")[:] = D
    s.frw_set(gs["phi_frw"], 0.0)
    s.frw_add_fluid(gs["rho_frw"], 0.0)
    s.clearSrc()
    oracle_mod.lib().or_static_add_src(s.h)
    assert digest(s.field("DIFFr_a")) == gs["sha256"]


def test_rk4_register_bitwise_vs_reference(oracle_mod):
    r = GOLD["rk4"]
    x0 = np.ascontiguousarray(r["x0"])
    n = x0.size
    traj = np.zeros((r["steps"], n))
    last = np.zeros((4, n))
    oracle_mod.lib().or_rk4_demo(oracle_mod.dptr(x0), n, r["h"], r["c0"], r["c1"], r["steps"],
                                 oracle_mod.dptr(traj), oracle_mod.dptr(last))
    assert [float.hex(float(v)) for v in traj.reshape(-1)] == r["traj"]
    assert [float.hex(float(v)) for v in last.reshape(-1)] == r["last_stage_a"]


def test_frw_bitwise_vs_reference(oracle_mod):
    fr = GOLD["frw"]
    s = oracle_mod.Sim(oracle_mod.make_params(12, order=4))
    s.frw_set(fr["phi0"], fr["K0"])
    for rho, w in zip(fr["rho"], fr["w"]):
        s.frw_add_fluid(rho, w)
    got = []
    for _ in range(fr["steps"]):
        for st in range(1, 5):
            oracle_mod.lib().or_frw_step_stage(s.h, st, fr["h"])
            got.extend(s.frw_get())
    assert [float.hex(float(v)) for v in got] == fr["stage_values"]


def test_rk4_known_answer(oracle_mod):
    """tests/rk4.cc:24-62 of the reference, restated on the oracle's registers:
    x' = exp(-x), x(1) = 0, dt = 0.1, 10000 steps, |x - ln t| < 1e-4."""
    x0 = np.zeros(1)
    traj = np.zeros((10000, 1))
    last = np.zeros((4, 1))
    oracle_mod.lib().or_rk4_demo(oracle_mod.dptr(x0), 1, 0.1, 0.0, 0.0, 10000,
                                 oracle_mod.dptr(traj), oracle_mod.dptr(last))
    assert abs(traj[-1, 0] - np.log(1.0 + 10000 * 0.1)) < 1e-4
    exe = os.path.join(oracle_mod.REF_DIR, "rk4_kat")
    if os.path.exists(exe):
        out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0, out.stdout + out.stderr
        # the reference prints its final value with 6 significant digits
        assert f"{traj[-1, 0]:.6g}" in out.stdout


def test_frw_known_answer_reference_binary(oracle_mod):
    exe = os.path.join(oracle_mod.REF_DIR, "frw_kat")
    if not os.path.exists(exe):
        pytest.skip("reference build not present (oracle/_ref)")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr


def test_frw_friedmann_invariant(oracle_mod):
    """tests/reference_frw.cc:12-51 restated on the oracle's FRW: four fluids
    (w = 0, 1/3, -1, -1/2), dt = 0.01 to t = 1, |K^2 - 24 pi rho| < dt^2."""
    s = oracle_mod.Sim(oracle_mod.make_params(12, order=4))
    H0 = 2.89440502
    s.frw_set(0.0, -3.0 * H0)
    for w in (0.0, 1.0 / 3.0, -1.0, -0.5):
        s.frw_add_fluid(0.25, w)
    dt, t = 0.01, 0.0
    while t < 1:
        for st in range(1, 5):
            oracle_mod.lib().or_frw_step_stage(s.h, st, dt)
        t += dt
    phi, K, rho, S = s.frw_get()
    assert abs(K * K - 24 * np.pi * rho) < dt * dt
    # matter-era closed form on its own: a = (1 + 3 t H0 / 2)^(2/3), phi = ln(a)/2
    s2 = oracle_mod.Sim(oracle_mod.make_params(12, order=4))
    rho0 = 3.0 / (8 * np.pi)
    s2.frw_set(0.0, -3.0)
    s2.frw_add_fluid(rho0, 0.0)
    for _ in range(200):
        for st in range(1, 5):
            oracle_mod.lib().or_frw_step_stage(s2.h, st, 0.005)
    a = (1 + 1.5 * 1.0) ** (2.0 / 3.0)
    phi, K, rho, S = s2.frw_get()
    assert abs(phi - 0.5 * np.log(a)) < 1e-9
    assert abs(K - (-3.0 / (1 + 1.5))) < 1e-9
    assert abs(rho - rho0 / a**3) < 1e-9


@pytest.mark.parametrize("order", [4, 6])
def test_stencils_live_vs_reference(oracle_mod, order):
    if not oracle_mod.ref_available(order):
        pytest.skip("reference build not present (oracle/_ref)")
    R = oracle_mod.ref_lib(order)
    nx, ny, nz, _ = oracle_mod.ref_dims(R)
    f = np.ascontiguousarray(np.random.default_rng(77).standard_normal((nx, ny, nz)))
    R.ref_set_steps(1.0 / nx, 0.1 / nx)
    d1 = np.zeros((3, nx, ny, nz))
    R.ref_derivatives(oracle_mod.dptr(f), oracle_mod.dptr(d1))
    _, m1, _, _ = oracle_stencils(oracle_mod, order, f, 0.0)
    assert np.array_equal(d1, m1)


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_upwind_bitwise_vs_reference(oracle_mod, order):
    """upwind_derivative (utils/math.h:1247-1252: the lopsided derivative plus
    its dissipation stencil, forward for a positive speed, backward
    otherwise) with a signed per-point speed, bitwise; the lopsided stencils
    are stubs returning 0 at orders 6 and 8 (math.h:539-549, 661-673)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    shape = tuple(GOLD["grid"])
    f = np.ascontiguousarray(field_input(shape, GOLD["seed_field"] + order))
    c = mg.upwind_speed(shape, order)
    s = oracle_mod.Sim(oracle_mod.make_params(shape, order=order))
    up = oracle_mod.upwind(s, f, c)
    assert digest(up) == GOLD["stencils"][str(order)]["upwind"]["sha256"]
    assert (np.abs(up).max() > 0) == (order in (2, 4))
    if oracle_mod.ref_available(order):
        R = oracle_mod.ref_lib(order)
        R.ref_set_steps(1.0 / shape[0], 0.1 / shape[0])
        ref = np.zeros((3,) + shape)
        R.ref_upwind(oracle_mod.dptr(f), oracle_mod.dptr(c), oracle_mod.dptr(ref))
        assert np.array_equal(ref.view(np.int64), up.view(np.int64))


@pytest.mark.parametrize("order", [2, 4])
def test_upwind_symbol(oracle_mod, order):
    """On a Fourier mode e^{i k x}, the upwind derivative is c * sigma(k) e^{ikx}
    with the lopsided stencil's symbol: order 4 forward
    (e^{3ih}/12 - e^{2ih}/2 + 3e^{ih}/2 - 5/6 - e^{-ih}/4)/dx, order 2 the
    one-sided second-order stencil minus dx/2 times the forward second
    difference; the backward symbols are the complex conjugates times -1."""
    n = 16
    s = oracle_mod.Sim(oracle_mod.make_params((n, 4, 4), order=order))
    dx = 1.0 / n
    kk = 3
    x = np.arange(n) * dx
    h = 2 * np.pi * kk * dx
    z = np.exp(1j * h)
    if order == 4:
        fw = (z**3 / 12 - z**2 / 2 + 1.5 * z - 5.0 / 6 - 1 / (4 * z)) / dx
    else:
        fw = (-0.5 * z**2 + 2 * z - 1.5) / dx - 0.5 * dx * (2 - 5 * z + 4 * z**2 - z**3) / dx**2
    for part in (np.cos, np.sin):
        f = np.ascontiguousarray(np.broadcast_to(part(2 * np.pi * kk * x)[:, None, None], (n, 4, 4)))
        for sign in (1.0, -1.0):
            c = np.full((n, 4, 4), 0.7 * sign)
            up = oracle_mod.upwind(s, f, c)[0][:, 0, 0]
            sym = fw if sign > 0 else -np.conj(fw)
            e = np.exp(1j * 2 * np.pi * kk * x)
            want = 0.7 * sign * (sym * e).real if part is np.cos else 0.7 * sign * (sym * e).imag
            np.testing.assert_allclose(up, want, rtol=0, atol=1e-11 / dx)
