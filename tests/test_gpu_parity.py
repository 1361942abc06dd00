"""GPU parity against the CPU oracle (FP64 drop-in and FP32 split).

Every case creates the oracle and the GPU context from identical data and
compares every field register.  Bars: FP64 state <= 1e-12 relative L2
(the drop-in's parity bar), single right-hand sides <= 1e-10 relative L2 (a RHS
carries the cancellation of the stencils; the state it feeds does not);
FP32 split <= 1e-5 relative L2 of the state.
"""
import sys

import numpy as np
import pytest

from common import FIELDS, rel_l2, setup_pair, smooth_random_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_rhs_parity_random_state(oracle_mod, gpu, order):
    shape = (16, 14, 18)
    _say(f"case rhs order={order} shape={shape}")
    st = smooth_random_state(shape, seed=order)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": order, "lapse": "Harmonic"}, state=st)
    o.stepInit(); o.clearSrc(); oracle_mod.lib().or_static_add_src(o.h); oracle_mod.lib().or_lambda_add_src(o.h)
    o.RKEvolve()
    g.stepInit()
    g.RKEvolve()
    for f in FIELDS:
        a_err = rel_l2(g.download(f + "_a"), o.field(f + "_a"))
        assert a_err <= 1e-14, (f, "_a", a_err)
        c_err = rel_l2(g.download(f + "_c"), o.field(f + "_c"))
        assert c_err <= 1e-10, (f, "_c", c_err)


@pytest.mark.parametrize("order,ko", [(2, 0.3), (8, 0.05)])
def test_rhs_parity_ko(oracle_mod, gpu, order, ko):
    shape = (16, 16, 20)
    _say(f"case ko order={order} ko={ko}")
    st = smooth_random_state(shape, seed=10 + order, kmax=3)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": order, "ko": ko}, state=st)
    o.stepInit(); o.clearSrc(); oracle_mod.lib().or_static_add_src(o.h)
    o.RKEvolve()
    g.stepInit()
    g.RKEvolve()
    for f in FIELDS:
        assert rel_l2(g.download(f + "_c"), o.field(f + "_c")) <= 1e-10, f


def test_method_sequence_parity(oracle_mod, gpu):
    """stepInit / RKEvolve / K1..K4Finalize, one step, every register."""
    shape = (16, 16, 16)
    _say("case method-sequence")
    st = smooth_random_state(shape, seed=3)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4}, state=st)
    L = oracle_mod.lib()
    o.stepInit()
    g.stepInit()
    for n in range(1, 5):
        o.clearSrc(); L.or_static_add_src(o.h); L.or_lambda_add_src(o.h)
        o.RKEvolve(); o.KFinalize(n)
        g.RKEvolve(); getattr(g, f"K{n}Finalize")()
        for f in FIELDS:
            for r in "pacf":
                err = rel_l2(g.download(f"{f}_{r}"), o.field(f"{f}_{r}"))
                assert err <= 1e-12, (n, f, r, err)
    assert np.allclose(g.frw_get(), o.frw_get(), rtol=0, atol=0)


@pytest.mark.parametrize("order,n,steps", [(4, 16, 10), (6, 16, 6)])
def test_fused_steps_random_state(oracle_mod, gpu, order, n, steps):
    _say(f"case fused order={order} n={n} steps={steps}")
    st = smooth_random_state((n, n, n), seed=5)
    o, g = setup_pair(oracle_mod, gpu, {"n": n, "order": order}, state=st)
    o.run_steps(steps)
    g.step(steps)
    for f in FIELDS:
        err = rel_l2(g.download(f + "_p"), o.field(f + "_p"))
        assert err <= 1e-12, (f, err)
    assert g.frw_get() == o.frw_get()


def test_config1_static_dust_50_steps(oracle_mod, gpu):
    """32^3 static dust, order 4, P(k)~k^-2, rms 1e-3, seed 9, 50 RK4 steps."""
    from cosmograph_amd.ics import static_dust_ics
    _say("case config1 32^3 x 50 steps")
    ics = static_dust_ics(32, -2.0, 1e-3, 9)
    o, g = setup_pair(oracle_mod, gpu, {"n": 32, "order": 4}, ics=ics)
    o.run_steps(50)
    g.step(50)
    worst = 0.0
    for f in FIELDS:
        ref = o.field(f + "_p")
        err = rel_l2(g.download(f + "_p"), ref)
        if np.any(ref != 0):
            worst = max(worst, err)
            assert err <= 1e-12, (f, err)
        else:
            assert err <= 1e-30, (f, err)
    _say(f"config1 worst rel L2 {worst:.3e}")


def test_fp32_split_random_state(oracle_mod, gpu):
    shape = (24, 24, 24)
    _say("case fp32-split")
    st = smooth_random_state(shape, seed=7, with_alpha=False)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4, "precision": "fp32_split"}, state=st)
    o.run_steps(8)
    g.step(8)
    for f in FIELDS:
        ref = o.field(f + "_p")
        err = rel_l2(g.download(f + "_p"), ref)
        assert err <= 1e-5, (f, err)


def test_autotuned_variant_parity(oracle_mod, gpu):
    """Whatever variant the autotuner picks must give the same evolution."""
    shape = (32, 32, 32)
    _say("case autotune")
    st = smooth_random_state(shape, seed=11, with_alpha=False)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4}, state=st)
    report = g.autotune(2)
    assert len(report) >= 1
    o.run_steps(3)
    g.step(3)
    for f in FIELDS:
        assert rel_l2(g.download(f + "_p"), o.field(f + "_p")) <= 1e-12, f


@pytest.mark.parametrize("tile,precision,tol", [((2, 8, 16), "fp64", 1e-12), ((4, 4, 16), "fp64", 1e-12),
                                                ((8, 8, 8), "fp32_split", 1e-5), ((4, 4, 16), "fp32_fix", 2e-5),
                                                ((4, 8, 8), "fp32_fix", 2e-5), ((4, 4, 16), "fp32_all", 1e-4)])
def test_tile_variants_parity(oracle_mod, gpu, tile, precision, tol):
    """Each tile against the oracle on a grid ragged along k (nz = 40), with
    LDS-DMA staging: in the 4-byte modes at order 6 that includes the two-field
    tiles of K and the Gamma^i."""
    shape = (24, 24, 40)
    _say(f"case tile {tile} {precision}")
    st = smooth_random_state(shape, seed=12)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 6, "precision": precision}, state=st,
                      gpu_kw={"tile": tile})
    assert g.kernel_info()["tile"] == tile
    o.run_steps(2)
    g.step(2)
    for f in FIELDS:
        assert rel_l2(g.download(f + "_p"), o.field(f + "_p")) <= tol, f


def _oracle_output_step(oracle_mod, o):
    """The reference's output step sees _a after initStaticStep (stepInit +
    sources, sims/static.cc:74-82)."""
    L = oracle_mod.lib()
    o.stepInit()
    o.clearSrc()
    L.or_static_add_src(o.h)
    L.or_lambda_add_src(o.h)
    return o.constraints()


def _check_stats(got, want, rel):
    for c in "HMGAS":
        g, w = np.array(got[c]), np.array(want[c])
        if abs(w[2]) < 1e-14:  # constraint satisfied to roundoff: compare magnitudes only
            assert abs(g[2]) < 1e-14, (c, g[2])
            continue
        scale = max(abs(w[2]), 1e-300)
        # mean / stdev / max of values, sum of scales, scaled statistics, L2
        assert abs(g[0] - w[0]) <= rel * scale, (c, "mean", g[0], w[0])
        for q in (1, 2, 3, 5, 6) + ((7,) if c in "HM" else ()):
            assert abs(g[q] - w[q]) <= rel * max(abs(w[q]), 1e-300), (c, q, g[q], w[q])
        if np.isfinite(w[4]):
            assert abs(g[4] - w[4]) <= rel * max(abs(w[6]), 1e-300), (c, "mean_scaled", g[4], w[4])


@pytest.mark.parametrize("order", [4, 6])
def test_constraints_vs_oracle(oracle_mod, gpu, order):
    shape = (20, 18, 24)
    _say(f"case constraints order={order}")
    st = smooth_random_state(shape, seed=21 + order)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": order}, state=st)
    want = _oracle_output_step(oracle_mod, o)
    got = g.setConstraintCalcs()
    _check_stats(got, want, 1e-10)
    np.testing.assert_allclose(g.download("ricci_a"), o.field("ricci_a"), rtol=1e-11, atol=1e-11 * np.abs(o.field("ricci_a")).max())


def test_constraints_after_steps_config1(oracle_mod, gpu):
    from cosmograph_amd.ics import static_dust_ics
    _say("case constraints config1 after 10 steps")
    ics = static_dust_ics(32, -2.0, 1e-3, 9)
    o, g = setup_pair(oracle_mod, gpu, {"n": 32, "order": 4}, ics=ics)
    o.run_steps(10)
    g.step(10)
    _check_stats(g.setConstraintCalcs(), _oracle_output_step(oracle_mod, o), 1e-10)


def test_constraints_fp32_split(oracle_mod, gpu):
    shape = (24, 24, 24)
    _say("case constraints fp32")
    st = smooth_random_state(shape, seed=31, with_alpha=False)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4, "precision": "fp32_split"}, state=st)
    want = _oracle_output_step(oracle_mod, o)
    got = g.setConstraintCalcs()
    for c in "HMG":
        for q in (2, 6) + ((7,) if c in "HM" else ()):
            assert abs(got[c][q] - want[c][q]) <= 1e-5 * abs(want[c][q]), (c, q)
    # algebraic constraints hold to the storage precision: tr A_ij of FP32
    # A_ij (|A| ~ 1e-3), det(gamma) - 1 of the FP64 metric
    assert got["A"][2] < 1e-8
    assert got["S"][2] < 1e-12
