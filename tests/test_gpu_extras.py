"""GPU parity of setExtraFieldData (bssn.cc:165-170), the ConformalFLRW lapse
that consumes K_avg every stage (BSSNGaugeHandler.cc:24-27), and numNaNs
(utils/math.h:1524-1541), against the CPU oracle.

Bars: FP64 state <= 1e-12 relative L2; the four averages <= 1e-12 relative
(they are sums over the grid taken in a different order); NaN counts exact.
"""
import math
import sys

import numpy as np
import pytest

from common import FIELDS, rel_l2, setup_pair, smooth_random_state

pytestmark = pytest.mark.gpu

EXTRA = ("K_min", "K_avg", "avg_vol", "rho_avg")


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check_extra(got, want, what, rtol=1e-12, scale=None):
    """Averages agree to rtol relative to the average, or to the size of the
    averaged values (`scale`) when the average itself cancels (a zero-mean
    perturbation such as DIFFK or DIFFr of constraint-solved initial data)."""
    for k in EXTRA:
        w = want[k]
        tol = rtol * max(abs(w), (scale or {}).get(k, 0.0), 1e-300)
        assert abs(got[k] - w) <= tol, (what, k, got[k], w)


def _check_fields(g, o, regs="p", tol=1e-12):
    for f in FIELDS:
        for r in regs:
            err = rel_l2(g.download(f"{f}_{r}"), o.field(f"{f}_{r}"))
            assert err <= tol, (f, r, err)


def test_extra_field_data_method_level(oracle_mod, gpu):
    """stepInit / RKEvolve / K{n}Finalize, ConformalFLRW lapse: the averages
    after every call and every register."""
    shape = (18, 14, 20)  # ragged against every tile shape
    _say(f"case extras method-level {shape}")
    st = smooth_random_state(shape, seed=41)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4, "lapse": "ConformalFLRW"}, state=st)
    L = oracle_mod.lib()
    for step in range(2):
        o.stepInit()
        g.stepInit()
        ex = g.extra_field_data()
        want = o.extra()
        # stepInit keeps rho_avg of the last stage (none before the first)
        _check_extra(ex, want, ("stepInit", step))
        for n in range(1, 5):
            o.clearSrc(); L.or_static_add_src(o.h); L.or_lambda_add_src(o.h)
            o.RKEvolve(); o.KFinalize(n)
            g.RKEvolve(); getattr(g, f"K{n}Finalize")()
            _check_extra(g.extra_field_data(), o.extra(), ("K", n, step))
        _check_fields(g, o, "pacf")


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32_split", 1e-5)])
def test_conformal_flrw_lapse_fused(oracle_mod, gpu, precision, tol):
    shape = (18, 14, 20)
    _say(f"case ConformalFLRW fused {precision}")
    st = smooth_random_state(shape, seed=43)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4, "lapse": "ConformalFLRW",
                                        "precision": precision}, state=st)
    o.run_steps(4)
    g.step(4)
    _check_fields(g, o, "p", tol)
    _check_extra(g.extra_field_data(), o.extra(), "after 4 steps", 1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("order,n", [(4, 16), (6, 24)])
def test_extra_field_data_fused_static(oracle_mod, gpu, order, n):
    """Static lapse: the averages of the last K4Finalize are formed on demand;
    a later upload must not change them (the reference's members keep the
    values of the last setExtraFieldData call)."""
    _say(f"case extras fused static order={order} n={n}")
    st = smooth_random_state((n, n, n), seed=44, with_alpha=False)
    o, g = setup_pair(oracle_mod, gpu, {"n": n, "order": order, "lam": 0.01}, state=st)
    o.run_steps(3)
    g.step(3)
    want = o.extra()
    # upload first, read afterwards: the values of the step must survive
    g.upload("DIFFK_p", g.download("DIFFK_p") + 1.0)
    _check_extra(g.extra_field_data(), want, "after upload")


def test_extra_field_data_config1(oracle_mod, gpu):
    from cosmograph_amd.ics import static_dust_ics
    _say("case extras config1 32^3 x 5 steps")
    ics = static_dust_ics(32, -2.0, 1e-3, 9)
    o, g = setup_pair(oracle_mod, gpu, {"n": 32, "order": 4}, ics=ics)
    for _ in range(5):
        o.run_steps(1)
        g.step(1)
        scale = {"K_avg": np.abs(o.field("DIFFK_p")).max(), "rho_avg": np.abs(o.field("DIFFr_a")).max()}
        _check_extra(g.extra_field_data(), o.extra(), "config1", scale=scale)
    # the expansion goal of CosmoSim::run (sims/sim.cc:114-121) reads avg_vol
    assert g.extra_field_data()["avg_vol"] > 0


def test_num_nans_and_nan_guard(oracle_mod, gpu):
    shape = (16, 16, 16)
    _say("case numNaNs")
    st = smooth_random_state(shape, seed=45)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4}, state=st)
    k = st["DIFFK"].copy()
    k[1, 2, 3] = np.nan
    k[4, 5, 6] = np.nan
    k[7, 8, 9] = 1e39  # rounds to inf in float: not a NaN for numNaNs
    g.upload("DIFFK_p", k)
    o.field("DIFFK_p")[:] = k
    assert g.numNaNs("DIFFK_p") == 2
    assert o.num_nans("DIFFK_p") == 2
    assert g.numNaNs("DIFFphi_p") == 0
    # the per-step guard of CosmoSim::runCommonStepTasks (sims/sim.cc:179-183)
    ph = st["DIFFphi"].copy()
    ph[0, 0, 0] = np.nan
    g.upload("DIFFphi_p", ph)
    with pytest.raises(gpu.CosmoGpuError) as ei:
        g.step(1)
    assert ei.value.code == gpu.ERR_NAN
    assert g.numNaNs("DIFFphi_p") > 0


@pytest.mark.parametrize("lapse", ["Anharmonic", "OnePlusLog", "ConformalSync", "AwAGaugeWave",
                                   "AwAShiftedWave", "DampedWave", "TestAijDriverLapse"])
def test_lapse_gauges_vs_oracle(oracle_mod, gpu, lapse):
    """The shift-free lapse gauges of BSSNGaugeHandler on a generic state
    (K_avg read from the device every stage for Anharmonic and OnePlusLog)."""
    shape = (16, 14, 18)
    _say(f"case lapse {lapse}")
    st = smooth_random_state(shape, seed=47)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4, "lapse": lapse}, state=st)
    if lapse in ("OnePlusLog", "DampedWave"):  # non-default gauge parameters
        o = oracle_mod.Sim(oracle_mod.make_params(shape, order=4, lapse=lapse, gd_c=0.7, dw_mu_l=0.3, dw_p=0.5))
        for f in FIELDS:
            o.field(f + "_p")[:] = st[f]
        o.field("DIFFD_a")[:] = st["DIFFD"]
        o.frw_set(0.0, -3.0)
        o.frw_add_fluid(3.0 / (8.0 * np.pi), 0.0)
        g.L.cg_set_gauge_params(g.h, 0.7, 0.3, 0.5)
    o.run_steps(3)
    g.step(3)
    _check_fields(g, o, "p", 1e-12)
