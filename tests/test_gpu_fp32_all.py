"""The all-FP32 storage mode (CG_FP32_ALL: every BSSN field stored, staged
and differentiated in FP32, the conformal metric included; RK4 accumulation,
FRW background and reductions in FP64) against the FP64 oracle, through the
C ABI.

Bars: smooth random data of amplitude 1e-3 over a few steps, relative L2
<= 2e-5 per field (FP32 rounding of stored fields whose second derivatives
enter the Ricci tensor); the fused path equals the method-level path to
FP32 rounding of the accumulator-free form (<= 5e-6); constraint norms
within 1e-4 of the oracle's.
"""
import sys

import numpy as np
import pytest

from common import FIELDS, rel_l2, setup_pair, smooth_random_state

pytestmark = pytest.mark.gpu

SHAPE = (20, 16, 24)  # nz % 4 == 0: the LDS-DMA kernels apply


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("order,lapse", [(4, "Static"), (6, "Static"), (4, "Harmonic")])
def test_fp32_all_steps_vs_oracle(oracle_mod, gpu, order, lapse):
    _say(f"case fp32_all steps order {order} lapse {lapse}")
    st = smooth_random_state(SHAPE, seed=21, with_alpha=lapse != "Static")
    o, g = setup_pair(oracle_mod, gpu, {"n": SHAPE, "order": order, "lapse": lapse, "precision": "fp32_all"},
                      state=st)
    o.run_steps(4)
    g.step(4)
    worst = 0.0
    for f in FIELDS:
        ref = o.field(f + "_p")
        err = rel_l2(g.download(f + "_p"), ref)
        if np.any(ref != 0):
            worst = max(worst, err)
            assert err <= 2e-5, (f, err)
        else:
            assert err == 0.0, (f, err)
    _say(f"  worst rel L2 {worst:.2e}")
    # the output step's state on the oracle side: stepInit and the sources
    L = o.L
    L.or_bssn_step_init(o.h)
    L.or_bssn_clear_src(o.h)
    L.or_static_add_src(o.h)
    L.or_lambda_add_src(o.h)
    c32, c64 = g.setConstraintCalcs(), o.constraints()
    for q in (2, 7):
        assert abs(c32["H"][q] / c64["H"][q] - 1) <= 1e-4, (q, c32["H"][q], c64["H"][q])


def test_fp32_all_method_level_stages(oracle_mod, gpu):
    _say("case fp32_all method level")
    st = smooth_random_state(SHAPE, seed=22)
    o, g = setup_pair(oracle_mod, gpu, {"n": SHAPE, "order": 4, "lapse": "Harmonic", "precision": "fp32_all"},
                      state=st)
    L = o.L
    o.stepInit()
    g.stepInit()
    for n in range(1, 5):
        o.clearSrc(); L.or_static_add_src(o.h); L.or_lambda_add_src(o.h)
        o.RKEvolve(); o.KFinalize(n)
        g.RKEvolve(); getattr(g, f"K{n}Finalize")()
        for f in FIELDS:
            for r in "pac":
                ref = o.field(f"{f}_{r}")
                assert rel_l2(g.download(f"{f}_{r}"), ref) <= 2e-5, (n, f, r)


def test_fp32_all_fused_equals_method_level(gpu):
    """Fused steps (DMA kernels, normalised copy written by stage 4) against
    the method-level sequence on the same data."""
    _say("case fp32_all fused vs method")
    st = smooth_random_state(SHAPE, seed=23, with_alpha=False)
    runs = []
    for fused in (True, False):
        g = gpu.BSSN(gpu.GpuParams(n=SHAPE, order=6, precision="fp32_all"))
        for f in FIELDS:
            g.upload(f + "_p", st[f])
        g.upload("DIFFD_a", st["DIFFD"])
        g.frw_set(0.0, -3.0)
        g.frw_add_fluid(3.0 / (8.0 * np.pi), 0.0)
        if fused:
            g.step(3)
        else:
            for _ in range(3):
                g.stepInit()
                for n in range(1, 5):
                    g.RKEvolve()
                    getattr(g, f"K{n}Finalize")()
        runs.append({f: g.download(f + "_p") for f in FIELDS})
    for f in FIELDS:
        assert rel_l2(runs[0][f], runs[1][f]) <= 5e-6, f


def test_fp32_all_device_ics(gpu):
    """cg_ics_static_dust writes the FP32 metric: DIFFphi agrees with the
    split context's FP64 DIFFphi to FP32 rounding."""
    from cosmograph_amd.ics import gaussian_modes
    _say("case fp32_all ICs")
    modes = gaussian_modes(32, -2.0, 4)
    out = {}
    for prec in ("fp32_all", "fp32_split"):
        g = gpu.BSSN(gpu.GpuParams(n=32, order=4, precision=prec))
        g.ics_static_dust(modes, 1e-3)
        out[prec] = (g.download("DIFFphi_p"), g.download("DIFFD_a"), g.frw_get())
    assert rel_l2(out["fp32_all"][0], out["fp32_split"][0]) <= 1e-7
    assert np.array_equal(out["fp32_all"][1], out["fp32_split"][1])
    assert out["fp32_all"][2] == out["fp32_split"][2]


def test_fp32_all_unsupported(gpu):
    g = gpu.BSSN(gpu.GpuParams(n=16, order=4, precision="fp32_all"))
    g.frw_set(0.0, 0.0)
    g.stepInit()
    with pytest.raises(gpu.CosmoGpuError):
        g.bardeen(0.0)
    with pytest.raises(gpu.CosmoGpuError):
        g.field_stats("DIFFK_a")
    with pytest.raises(gpu.CosmoGpuError):
        gpu.Batch([gpu.GpuParams(n=16, order=4, precision="fp32_all")])
