"""The fixed-point metric mode (CG_FP32_FIX: DIFFgamma_ij and DIFFphi as
32-bit words with one power-of-two scale per field, differentiated by FP32
sums of exact integer differences; every other BSSN field in FP32; RK4
accumulation, FRW background and reductions in FP64) against the FP64
oracle, through the C ABI.

Bars: smooth random data of amplitude 1e-3 over a few steps, relative L2
<= 2e-6 for the metric fields (the field's maximum in 29-bit words) and,
for the FP32 fields, <= 2e-5 or no more than twice the error of the split
mode (FP64 metric, the same FP32 fields) on the same data;
constraint norms within 1e-4 of the oracle's; an upload reads back within
half a word; exponent changes (a field's scale raised or lowered between
steps) keep the same bars.
"""
import sys

import numpy as np
import pytest

from common import FIELDS, rel_l2, setup_pair, smooth_random_state

pytestmark = pytest.mark.gpu

SHAPE = (20, 16, 24)  # nz % 4 == 0: the LDS-DMA kernels apply
METRIC = FIELDS[:7]


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(o, g, what, bar=2e-5, metric_bar=2e-6, split=None):
    """split: the split mode's relative errors on the same data (FP32 fields)"""
    worst = 0.0
    for f in FIELDS:
        ref = o.field(f + "_p")
        err = rel_l2(g.download(f + "_p"), ref)
        if np.any(ref != 0):
            worst = max(worst, err)
            lim = metric_bar if f in METRIC else max(bar, 2.0 * split[f]) if split else bar
            assert err <= lim, (what, f, err, lim)
        else:
            assert err == 0.0, (what, f, err)
    return worst


@pytest.mark.parametrize("order,lapse", [(4, "Static"), (6, "Static"), (4, "Harmonic"), (6, "OnePlusLog")])
def test_fp32_fix_steps_vs_oracle(oracle_mod, gpu, order, lapse):
    _say(f"case fp32_fix steps order {order} lapse {lapse}")
    st = smooth_random_state(SHAPE, seed=31, with_alpha=lapse != "Static")
    o, g = setup_pair(oracle_mod, gpu, {"n": SHAPE, "order": order, "lapse": lapse, "precision": "fp32_fix"},
                      state=st)
    _, gs = setup_pair(oracle_mod, gpu, {"n": SHAPE, "order": order, "lapse": lapse, "precision": "fp32_split"},
                       state=st)
    o.run_steps(5)
    g.step(5)
    gs.step(5)
    split = {f: rel_l2(gs.download(f + "_p"), o.field(f + "_p")) for f in FIELDS}
    _say(f"  worst rel L2 {_check(o, g, 'steps', split=split):.2e} (split mode {max(split.values()):.2e})")
    # the output step's state on the oracle side: stepInit and the sources
    L = o.L
    L.or_bssn_step_init(o.h)
    L.or_bssn_clear_src(o.h)
    L.or_static_add_src(o.h)
    L.or_lambda_add_src(o.h)
    cq, c64 = g.setConstraintCalcs(), o.constraints()
    for q in (2, 7):
        assert abs(cq["H"][q] / c64["H"][q] - 1) <= 1e-4, (q, cq["H"][q], c64["H"][q])


def test_fp32_fix_upload_round_trip(gpu):
    _say("case fp32_fix upload round trip")
    st = smooth_random_state(SHAPE, seed=32)
    g = gpu.BSSN(gpu.GpuParams(n=SHAPE, order=4, precision="fp32_fix"))
    g.frw_set(0.0, -3.0)
    for f in METRIC:
        x = st[f]
        g.upload(f + "_p", x)
        # exact to half a word: the largest |x| is at least 2^28 words
        assert np.abs(g.download(f + "_p") - x).max() <= np.abs(x).max() * 2.0 ** -28, f
    # stepInit copies _p to _a (DIFFphi unchanged); a 1000x larger upload to
    # DIFFphi_p raises the field's scale and re-expresses _a in it
    g.stepInit()
    x = st["DIFFphi"]
    g.upload("DIFFphi_p", 1000.0 * x)
    assert np.abs(g.download("DIFFphi_a") - x).max() <= 1000.0 * np.abs(x).max() * 2.0 ** -28
    assert np.abs(g.download("DIFFphi_p") - 1000.0 * x).max() <= 1000.0 * np.abs(x).max() * 2.0 ** -28


def test_fp32_fix_exponent_changes_vs_oracle(oracle_mod, gpu):
    """Between calls, a metric field is scaled up 20x (its exponent rises)
    and another down 200x (the next step's check lowers it): the steps keep
    the oracle's values to the same bars."""
    _say("case fp32_fix exponent changes")
    st = smooth_random_state(SHAPE, seed=33, with_alpha=False)
    o, g = setup_pair(oracle_mod, gpu, {"n": SHAPE, "order": 4, "lapse": "Static", "precision": "fp32_fix"},
                      state=st)
    o.run_steps(2)
    g.step(2)
    _check(o, g, "before")
    for f, s in (("DIFFgamma12", 20.0), ("DIFFgamma33", 1.0 / 200.0)):
        x = o.field(f + "_p") * s
        o.field(f + "_p")[:] = x
        g.upload(f + "_p", x)
    o.run_steps(5)
    g.step(5)
    _check(o, g, "after")


def test_fp32_fix_normalisation_sets_the_scale(oracle_mod, gpu):
    """A large off-diagonal component (0.5 here) makes the unit-determinant
    normalisation shift the diagonal components far beyond their own size;
    the exponent check covers the normalised values, so nothing saturates
    and the steps follow the oracle (bars of the words' rounding on the larger
    normalised range: 1e-5 on every field)."""
    _say("case fp32_fix normalisation range")
    st = smooth_random_state(SHAPE, seed=34, with_alpha=False)
    st["DIFFgamma12"] = st["DIFFgamma12"] * (0.5 / np.abs(st["DIFFgamma12"]).max())
    o, g = setup_pair(oracle_mod, gpu, {"n": SHAPE, "order": 4, "lapse": "Static", "precision": "fp32_fix"},
                      state=st)
    o.run_steps(3)
    g.step(3)
    _check(o, g, "normalised range", bar=1e-5, metric_bar=1e-5)


def test_fp32_fix_device_ics(gpu):
    """cg_ics_static_dust writes the fixed-point DIFFphi: it agrees with the
    split context's FP64 DIFFphi to half a word."""
    from cosmograph_amd.ics import gaussian_modes
    _say("case fp32_fix ICs")
    modes = gaussian_modes(32, -2.0, 4)
    out = {}
    for prec in ("fp32_fix", "fp32_split"):
        g = gpu.BSSN(gpu.GpuParams(n=32, order=4, precision=prec))
        g.ics_static_dust(modes, 1e-3)
        out[prec] = (g.download("DIFFphi_p"), g.download("DIFFD_a"), g.frw_get())
    ref = out["fp32_split"][0]
    assert np.abs(out["fp32_fix"][0] - ref).max() <= np.abs(ref).max() * 2.0 ** -28
    assert np.array_equal(out["fp32_fix"][1], out["fp32_split"][1])
    assert out["fp32_fix"][2] == out["fp32_split"][2]


def test_fp32_fix_unsupported(gpu):
    _say("case fp32_fix unsupported calls")
    g = gpu.BSSN(gpu.GpuParams(n=16, order=4, precision="fp32_fix"))
    g.frw_set(0.0, 0.0)
    g.stepInit()
    with pytest.raises(gpu.CosmoGpuError):
        g.RKEvolve()
    with pytest.raises(gpu.CosmoGpuError):
        g.K1Finalize()
    with pytest.raises(gpu.CosmoGpuError):
        g.bardeen(0.0)
    with pytest.raises(gpu.CosmoGpuError):
        g.strip("DIFFphi_p")
    with pytest.raises(gpu.CosmoGpuError):
        g.power_spectrum("DIFFgamma11_p")
    assert g.numNaNs("DIFFphi_p") == 0
    g.strip("DIFFK_p")  # the FP32 fields read as usual
    with pytest.raises(gpu.CosmoGpuError):
        gpu.Batch([gpu.GpuParams(n=16, order=4, precision="fp32_fix")])
