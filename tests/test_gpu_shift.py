"""GPU parity of the shift build (USE_BSSN_SHIFT, cosmo_macros.h:43-44) and
of the TestKDriverLapse against the CPU oracle, through the C ABI.

The shift build adds beta^i, expN (and auxB^i with the gamma driver) to the
evolved fields and the shift terms to every equation (bssn_macros.h:292-385,
bssn.cc:786-891), with the advection taken by the upwind stencils of
utils/math.h:1220-1252.  The oracle side is pinned by tests/test_oracle_shift.py
(exact gauge-wave solutions, ADM with Lie derivatives, bitwise upwind
stencils).  Data: smooth random periodic fields with a shift of either sign
(both upwind branches), on a grid ragged against the 4x8x8 tile.

Bars: FP64 <= 1e-12 relative L2 on every register of every field (absolute
L2 for identically-zero fields); FP32 split (A_ij, K, Gamma^i, alpha, D in
FP32) <= 1e-5 after three steps.
"""
import sys

import numpy as np
import pytest

from common import FIELDS, rel_l2, smooth_random_state

pytestmark = pytest.mark.gpu

SHAPE = (18, 14, 20)
X_ALL = ["beta1", "beta2", "beta3", "expN", "auxB1", "auxB2", "auxB3"]


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def x_fields(shift):
    return [] if shift == "none" else X_ALL if shift == "GammaDriver" else X_ALL[:4]


def shift_state(shape, seed, beta_amp=0.3):
    st = smooth_random_state(shape, seed=seed, amp=1e-3)
    extra = smooth_random_state(shape, seed=seed + 1000, amp=1.0)
    for i, f in enumerate(X_ALL):
        src = extra[FIELDS[i]]
        st[f] = (beta_amp if f.startswith("beta") else 1e-3) * src / np.abs(src).max()
    return st


def make_pair(po, cg, shape, order, shift, lapse="Static", precision="fp64", ko=0.0, seed=5, **gauge):
    gk = dict(gauge_wave_dir=2, gd_eta=0.4, dw_mu_s=0.3, dw_mu_l=0.2, dw_p=0.5, k_driver_coeff=0.07, gd_c=1.3)
    gk.update(gauge)
    o = po.Sim(po.make_params(shape, order=order, ko=ko, lapse=lapse, shift=shift, matter="static",
                              gauge_wave_dir=gk["gauge_wave_dir"], gd_eta=gk["gd_eta"], dw_mu_s=gk["dw_mu_s"],
                              dw_mu_l=gk["dw_mu_l"], dw_p=gk["dw_p"], k_driver_coeff=gk["k_driver_coeff"],
                              gd_c=gk["gd_c"]))
    g = cg.BSSN(cg.GpuParams(n=shape, order=order, precision=precision, ko_coeff=ko, lapse=lapse, shift=shift,
                             matter="static", **gk))
    st = shift_state(shape, seed)
    for f in FIELDS + x_fields(shift):
        o.field(f + "_p")[:] = st[f]
        g.upload(f + "_p", st[f])
    o.field("DIFFD_a")[:] = st["DIFFD"]
    g.upload("DIFFD_a", st["DIFFD"])
    for s in (o, g):
        s.frw_set(0.0, -3.0)
        s.frw_add_fluid(3.0 / (8.0 * np.pi), 0.0)
    return o, g


def oracle_rk_evolve(o):
    L = o.L
    o.clearSrc()
    L.or_static_add_src(o.h)
    L.or_lambda_add_src(o.h)
    o.RKEvolve()


def check(g, o, names, regs, tol, what):
    for f in names:
        for r in regs:
            err = rel_l2(g.download(f"{f}_{r}"), o.field(f"{f}_{r}"))
            assert err <= tol, (what, f, r, err)


CASES = [("Static", 4, "OnePlusLog"), ("GammaDriver", 4, "Harmonic"), ("DampedWave", 4, "DampedWave"),
         ("AwAShiftedWave", 4, "AwAShiftedWave"), ("AijDriverShift", 4, "Static"), ("DampedWave", 2, "OnePlusLog"),
         ("GammaDriver", 6, "Static"), ("AwAShiftedWave", 8, "AwAGaugeWave"), ("DampedWave", 6, "OnePlusLog"),
         ("AijDriverShift", 2, "TestKDriverLapse"), ("Static", 6, "TestKDriverLapse")]


@pytest.mark.parametrize("shift,order,lapse", CASES)
def test_shift_rhs_parity(oracle_mod, gpu, shift, order, lapse):
    """RKEvolve (stage kernel + gauge pass) writes every field's _c and the
    shift registers' _c as the oracle's RKEvolve."""
    _say(f"case shift rhs {shift} order {order} lapse {lapse} {SHAPE}")
    o, g = make_pair(oracle_mod, gpu, SHAPE, order, shift, lapse)
    o.stepInit()
    g.stepInit()
    oracle_rk_evolve(o)
    g.RKEvolve()
    check(g, o, FIELDS + x_fields(shift), "c", 1e-12, "rhs")


@pytest.mark.parametrize("shift,order,lapse,ko", [("DampedWave", 4, "OnePlusLog", 0.0),
                                                  ("GammaDriver", 2, "Harmonic", 0.05),
                                                  ("AijDriverShift", 8, "Static", 0.05)])
def test_shift_steps_parity(oracle_mod, gpu, shift, order, lapse, ko):
    """Three whole steps (cg_rk4_steps: the method sequence of
    sims/static.cc:74-170 with the gauge pass), KO on the shift as well at
    orders 2 and 8: every register of every field."""
    _say(f"case shift steps {shift} order {order} lapse {lapse} ko {ko}")
    o, g = make_pair(oracle_mod, gpu, SHAPE, order, shift, lapse, ko=ko)
    o.run_steps(3)
    g.step(3)
    check(g, o, FIELDS + x_fields(shift), "pacf", 1e-12, "steps")
    assert g.frw_get() == o.frw_get()


def test_shift_method_level_stages(oracle_mod, gpu):
    """stepInit, then RKEvolve + K{n}Finalize one by one: every register after
    every stage (the shift registers rotate with the others)."""
    _say("case shift method-level stages")
    o, g = make_pair(oracle_mod, gpu, SHAPE, 4, "GammaDriver", "OnePlusLog")
    o.stepInit()
    g.stepInit()
    for n in range(1, 5):
        oracle_rk_evolve(o)
        o.KFinalize(n)
        g.RKEvolve()
        getattr(g, f"K{n}Finalize")()
        check(g, o, FIELDS + X_ALL, "pacf", 1e-12, ("stage", n))


def test_k_driver_lapse_parity(oracle_mod, gpu):
    """TestKDriverLapse without the shift build: the lapse reads the complete
    d_t DIFFK and the K_avg / rho_avg of setExtraFieldData."""
    _say("case TestKDriverLapse")
    o, g = make_pair(oracle_mod, gpu, SHAPE, 4, "none", "TestKDriverLapse")
    o.run_steps(3)
    g.step(3)
    check(g, o, FIELDS, "p", 1e-12, "k-driver")
    assert rel_l2(g.download("DIFFalpha_p"), o.field("DIFFalpha_p")) <= 1e-12
    assert np.abs(o.field("DIFFalpha_p")).max() > 1e-4


def test_shift_fp32_split(oracle_mod, gpu):
    """FP32-split storage of A_ij, K, Gamma^i, alpha with the shift build
    (the shift registers and the metric stay FP64)."""
    _say("case shift fp32 split")
    o, g = make_pair(oracle_mod, gpu, SHAPE, 4, "DampedWave", "DampedWave", precision="fp32_split")
    o.run_steps(3)
    g.step(3)
    check(g, o, FIELDS + X_ALL[:4], "p", 1e-5, "fp32")


def test_shift_rejected_where_unsupported(gpu):
    """The shift build runs with static dust or vacuum on single grids."""
    with pytest.raises(gpu.CosmoGpuError):
        gpu.BSSN(gpu.GpuParams(n=16, order=4, matter="dust", shift="Static"))
    with pytest.raises(gpu.CosmoGpuError):
        gpu.Batch([gpu.GpuParams(n=16, order=4, shift="GammaDriver")])
