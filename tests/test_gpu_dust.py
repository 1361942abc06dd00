"""GPU parity of the dust fluid (components/dust_fluid/dust.cc, the DustSim
sequence sims/dust.cc:53-169) against the CPU oracle: the metric and the
four fluid registers after fused steps and after every method-level call,
the constraints with the fluid's sources, and the averages.  FP64 bar:
<= 1e-12 relative L2 on every register; FP32 split (dust registers, fluxes,
A_ij, K, Gamma^i in FP32, the metric in FP64): <= 1e-5."""
import sys

import numpy as np
import pytest

from common import FIELDS, rel_l2, setup_pair, smooth_random_state

pytestmark = pytest.mark.gpu

DUST = ["dust_D", "dust_S1", "dust_S2", "dust_S3"]
EMPTY_FRW = {"phi": 0.0, "K": 0.0, "fluids": []}
FULL_FRW = {"phi": 0.0, "K": -3.0, "fluids": [(3.0 / (8.0 * np.pi), 0.0)]}


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dust_pair(oracle_mod, gpu, shape, order=4, frw=EMPTY_FRW, lam=0.0, seed=61, lapse="Static", precision="fp64"):
    st = smooth_random_state(shape, seed=seed, with_alpha=(lapse != "Static"))
    if not frw["fluids"]:
        st["DIFFK"] = st["DIFFK"] - 0.5  # an expanding box without a reference background
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": order, "matter": "dust", "lam": lam,
                                        "lapse": lapse, "frw": frw, "precision": precision}, state=st)
    rng = np.random.default_rng(seed + 1)
    x = [np.arange(n) / n for n in shape]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    D = 0.05 * (1.0 + 0.02 * np.cos(2 * np.pi * (X + Y)) + 0.01 * np.sin(2 * np.pi * (Z - 2 * X)))
    S = [1e-3 * np.sin(2 * np.pi * (X + rng.uniform()) + 2 * np.pi * k * Z) for k in range(3)]
    for name, arr in zip(DUST, [D] + S):
        o.field(name + "_p")[:] = arr
        g.upload(name + "_p", arr)
    return o, g


def check(g, o, regs="p", tol=1e-12, names=None):
    for f in (names or FIELDS + DUST):
        for r in regs:
            ref = o.field(f"{f}_{r}")
            err = rel_l2(g.download(f"{f}_{r}"), ref)
            assert err <= (tol if np.any(ref != 0) else 1e-30), (f, r, err)


@pytest.mark.parametrize("order,frw,lam", [(4, "empty", 0.0), (6, "empty", 0.02), (4, "full", 0.0), (6, "full", 0.0)])
def test_dust_fused_steps(oracle_mod, gpu, order, frw, lam):
    """Three fused steps in one call.  At order 6 the LDS-DMA kernels run:
    stages 4 of the first two steps write the normalised copy of the new
    state, from which the next step's stage 1, its fluxes and stage 2's
    lagged sources read the metric."""
    shape = (18, 16, 20)
    _say(f"case dust fused order={order} frw={frw} lam={lam}")
    o, g = dust_pair(oracle_mod, gpu, shape, order, EMPTY_FRW if frw == "empty" else FULL_FRW, lam)
    o.run_steps(3)
    g.step(3)
    check(g, o)
    assert g.frw_get() == o.frw_get()


def test_dust_method_sequence(oracle_mod, gpu):
    """stepInit / RKEvolve / K{n}Finalize of BSSN and the fluid together, the
    sources rebuilt (with the lag) before every RKEvolve, as runDustStep does."""
    shape = (16, 16, 16)
    _say("case dust method-level")
    o, g = dust_pair(oracle_mod, gpu, shape, 4, EMPTY_FRW, 0.01)
    L = oracle_mod.lib()
    for step in range(2):
        L.or_bssn_step_init(o.h); L.or_dust_step_init(o.h)
        g.stepInit()
        for n in range(1, 5):
            L.or_bssn_clear_src(o.h); L.or_dust_add_src(o.h); L.or_lambda_add_src(o.h)
            L.or_bssn_rk_evolve(o.h); L.or_dust_rk_evolve(o.h)
            g.RKEvolve()
            check(g, o, "c", 1e-10)  # right-hand sides
            L.or_bssn_k_finalize(o.h, n); L.or_dust_k_finalize(o.h, n)
            getattr(g, f"K{n}Finalize")()
            check(g, o, "pacf")
        ex, want = g.extra_field_data(), o.extra()
        for k in ("K_min", "K_avg", "avg_vol", "rho_avg"):
            assert abs(ex[k] - want[k]) <= 1e-12 * abs(want[k]), (step, k)


def test_dust_constraints_and_averages(oracle_mod, gpu):
    shape = (20, 18, 24)
    _say("case dust constraints")
    o, g = dust_pair(oracle_mod, gpu, shape, 4, FULL_FRW, 0.0)
    o.run_steps(2)
    g.step(2)
    # averages of the last K4Finalize (DIFFr of stage 4, lagged)
    ex, want = g.extra_field_data(), o.extra()
    for k in ("K_min", "K_avg", "avg_vol"):
        assert abs(ex[k] - want[k]) <= 1e-12 * abs(want[k]), k
    scale = np.abs(o.field("DIFFr_a")).max()
    assert abs(ex["rho_avg"] - want["rho_avg"]) <= 1e-12 * max(abs(want["rho_avg"]), scale)
    np.testing.assert_allclose(g.download("DIFFr_a"), o.field("DIFFr_a"), rtol=1e-11, atol=1e-12 * scale)
    # the output step sees initDustStep's sources (sims/dust.cc:53-63)
    L = oracle_mod.lib()
    L.or_bssn_step_init(o.h); L.or_dust_step_init(o.h)
    L.or_bssn_clear_src(o.h); L.or_dust_add_src(o.h); L.or_lambda_add_src(o.h)
    want = o.constraints()
    got = g.setConstraintCalcs()
    for c in "HM":
        for q in (0, 2, 6, 7):
            w = want[c][q]
            assert abs(got[c][q] - w) <= 1e-10 * max(abs(w), abs(want[c][2])), (c, q, got[c][q], w)


@pytest.mark.parametrize("prec", ["fp32_split", "fp32_fix"])
@pytest.mark.parametrize("order,frw,lam", [(4, "empty", 0.0), (6, "empty", 0.02), (6, "full", 0.0)])
def test_dust_fp32_split(oracle_mod, gpu, order, frw, lam, prec):
    """The 4-byte modes: FP64 conformal metric (fp32_split) and the metric
    as 32-bit fixed point (fp32_fix), the dust registers and fluxes FP32."""
    shape = (18, 16, 20)
    _say(f"case dust {prec} order={order} frw={frw} lam={lam}")
    o, g = dust_pair(oracle_mod, gpu, shape, order, EMPTY_FRW if frw == "empty" else FULL_FRW, lam,
                     precision=prec)
    o.run_steps(3)
    g.step(3)
    check(g, o, tol=1e-5)
    want = o.field("DIFFr_a")
    assert rel_l2(g.download("DIFFr_a"), want) <= 1e-5


def test_dust_autotune_leaves_the_state(oracle_mod, gpu):
    """The autotuner's timing launches (stage-2 shape into dead registers)
    neither change the evolved state nor the kept source DIFFr_a."""
    shape = (16, 16, 20)
    _say("case dust autotune")
    o, a = dust_pair(oracle_mod, gpu, shape, 4, FULL_FRW, 0.0)
    _, b = dust_pair(oracle_mod, gpu, shape, 4, FULL_FRW, 0.0)
    a.step(1)
    b.step(1)
    r0 = b.download("DIFFr_a")
    b.autotune(2)
    assert np.array_equal(b.download("DIFFr_a"), r0)
    a.step(2)
    b.step(2)
    for f in FIELDS + DUST:
        assert np.array_equal(a.download(f + "_p"), b.download(f + "_p")), f


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32_fix", 1e-5)])
def test_dust_reference_rk4_form_vs_oracle(oracle_mod, gpu, prec, tol):
    """Fused dust steps with the reference's _f accumulator (the normalised
    copy then feeds stage 1 through slot Y, and stage 2's lagged metric
    comes from there) against the oracle."""
    shape = (16, 16, 40)
    _say(f"case dust reference RK4 form {prec}")
    o, g = dust_pair(oracle_mod, gpu, shape, 6, FULL_FRW, 0.0, precision=prec)
    g.set_fused_accumulator(True)
    o.run_steps(3)
    g.step(3)
    check(g, o, tol=tol)


@pytest.mark.parametrize("prec,lapse", [("fp32_fix", "OnePlusLog"), ("fp64", "Harmonic")])
def test_dust_live_lapse_vs_oracle(oracle_mod, gpu, prec, lapse):
    """A live lapse with the dust fluid (the kernels that stage and evolve
    DIFFalpha), order 6 on a grid with LDS-DMA staging: the fixed-point
    kernels with their epilogue prefetch, and FP64."""
    shape = (16, 16, 40)
    _say(f"case dust live lapse {prec} {lapse}")
    o, g = dust_pair(oracle_mod, gpu, shape, 6, FULL_FRW, 0.0, lapse=lapse, precision=prec)
    o.run_steps(3)
    g.step(3)
    check(g, o, tol=1e-5 if prec != "fp64" else 1e-12, names=FIELDS + DUST)


@pytest.mark.parametrize("prec", ["fp32_fix", "fp32_split"])
def test_dust_autotuned_vs_oracle(oracle_mod, gpu, prec):
    """The autotuner's pick (any tile, register budget and staging path that
    fits) steps the 4-byte dust modes to the oracle's bars."""
    shape = (16, 16, 40)
    _say(f"case dust autotuned {prec}")
    o, g = dust_pair(oracle_mod, gpu, shape, 6, FULL_FRW, 0.0, precision=prec)
    g.autotune(1)
    o.run_steps(3)
    g.step(3)
    check(g, o, tol=1e-5)


def _dust_ics(gpu, n, prec, lam, bg, modes):
    g = gpu.BSSN(gpu.GpuParams(n=n, order=6, precision=prec, matter="dust", lambda_=lam))
    g.set_dust_background(bg)
    g.ics_static_dust(modes, 3e-2)
    return g


def _totals(g):
    phiF, KF = g.frw_get()[:2]
    out = {f: g.download(f + "_p") for f in FIELDS + DUST}
    out["DIFFphi"] = out["DIFFphi"] + phiF
    out["DIFFK"] = out["DIFFK"] + KF
    return out


def test_dust_background_is_the_same_physics(gpu):
    """Dust as perturbations of the FRW fluids (mean dust w = 0, Lambda
    w = -1) evolves the same full fields as the reference's form with an empty
    FRW reference; FP32 split needs the perturbation form and matches FP64."""
    from cosmograph_amd.ics import RHO_BAR, gaussian_modes
    n, lam = 32, 0.7 * 3.0 / (8.0 * np.pi)
    _say(f"case dust background n={n}")
    modes = gaussian_modes(n, -3.0, 11)
    ref = _dust_ics(gpu, n, "fp64", lam, False, modes)
    pert = _dust_ics(gpu, n, "fp64", lam, True, modes)
    assert np.all(pert.download("DIFFK_p") == 0.0)
    assert abs(pert.frw_get()[2] - (RHO_BAR + lam)) <= 1e-15
    f32 = _dust_ics(gpu, n, "fp32_split", lam, True, modes)
    fix = _dust_ics(gpu, n, "fp32_fix", lam, True, modes)
    for g in (ref, pert, f32, fix):
        g.step(3)
    a, b, c, d = _totals(ref), _totals(pert), _totals(f32), _totals(fix)
    for f in FIELDS + DUST:
        if not np.any(a[f]):
            assert not np.any(b[f]) and not np.any(c[f]) and not np.any(d[f]), f
            continue
        assert rel_l2(b[f], a[f]) <= 1e-10, (f, rel_l2(b[f], a[f]))
        assert rel_l2(c[f], b[f]) <= 1e-5, (f, rel_l2(c[f], b[f]))
        assert rel_l2(d[f], b[f]) <= 1e-5, (f, rel_l2(d[f], b[f]))
