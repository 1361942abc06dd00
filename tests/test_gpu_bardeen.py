"""GPU parity of the Bardeen / SVT diagnostics (cg_bardeen: Bardeen::
setPotentials and getSVTViolations, components/bssn/bardeen.cc:22-467)
against the oracle's restatement (oracle/bardeen.py, pinned by
tests/test_oracle_bardeen.py), through the C ABI.

The reference computes them only without an FRW reference (bardeen.cc:26-31),
as in its dust runs: the data are smooth random perturbations around flat
space with static dust, a live lapse and, in one case, the shift build.
Inputs to both sides are the same state; the oracle's right-hand side comes
from the C oracle's RKEvolve.

Bars: FP64 <= 1e-10 relative L2 per potential (inverse Laplacians through
two different FFTs, hipFFT and numpy's, plus the 1e-13 right-hand-side
differences) and 1e-9 relative on each of the 17 violations; FP32 split
1e-4 (the stored A_ij, K, alpha enter d^2 h / dt^2).
"""
import sys

import numpy as np
import pytest

import bardeen as bd
from common import FIELDS, rel_l2, smooth_random_state

pytestmark = pytest.mark.gpu

NAMES = ["Phi", "Psi", "A", "dt_A", "d2t_A", "B", "dt_B", "d2t_B", "E", "F", "G1", "G2", "G3", "C1", "C2", "C3",
         "dt_C1", "dt_C2", "dt_C3", "Vmag", "D11", "D12", "D13", "D22", "D23", "D33"]
SHAPE = (16, 18, 20)


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_pair(po, cg, order, lapse, shift="none", precision="fp64", seed=11):
    dx = 1.0 / SHAPE[0]
    o = po.Sim(po.make_params(SHAPE, order=order, dx=dx, lapse=lapse, shift=shift, matter="static",
                              dw_mu_s=0.2, k_driver_coeff=0.05))
    g = cg.BSSN(cg.GpuParams(n=SHAPE, order=order, dx=dx, precision=precision, lapse=lapse, shift=shift,
                             matter="static", dw_mu_s=0.2, k_driver_coeff=0.05))
    st = smooth_random_state(SHAPE, seed=seed, amp=3e-3)
    st["DIFFalpha"] *= 10
    # a non-zero mean K, as in an expanding dust run: violations 9 and 12
    # divide by the (conformal / coordinate) mean of DIFFK, which a
    # zero-mean field leaves at round-off
    st["DIFFK"] += -0.05
    names = list(FIELDS)
    if shift != "none":
        for m in range(3):
            st[f"beta{m + 1}"] = 0.05 * smooth_random_state(SHAPE, seed=seed + 10 + m, amp=1.0)["DIFFK"]
            names.append(f"beta{m + 1}")
        st["expN"] = np.zeros(SHAPE)
        names.append("expN")
    for f in names:
        o.field(f + "_p")[:] = st[f]
        g.upload(f + "_p", st[f])
    D = 0.05 + 0.01 * st["DIFFD"] / np.abs(st["DIFFD"]).max()
    o.field("DIFFD_a")[:] = D
    g.upload("DIFFD_a", D)
    return o, g, names


def oracle_bardeen(o, names, order, t, use_ml, OL):
    L = o.L
    o.stepInit()
    o.clearSrc()
    L.or_static_add_src(o.h)
    L.or_lambda_add_src(o.h)
    o.RKEvolve()
    a_reg = {f: o.field(f + "_a").copy() for f in names}
    c_reg = {f: o.field(f + "_c").copy() for f in names}
    dx = o.p.dx
    return bd.set_potentials(a_reg, c_reg, t, order, dx, o.p.nx * dx, use_ml=use_ml, omega_L=OL)


def compare(g, want_f, want_v, got_v, tol_f, tol_v):
    for name in NAMES:
        got = g.download("Bardeen_" + name)
        w = want_f[name]
        err = rel_l2(got, w)
        assert err <= tol_f, (name, err, float(np.abs(w).max()))
    for q in range(17):
        w, v = want_v[q], got_v[q]
        if not np.isfinite(w):
            assert v == w or (np.isnan(v) and np.isnan(w)), (q, v, w)
            continue
        assert abs(v - w) <= tol_v * max(abs(w), 1e-300), (q, v, w)


@pytest.mark.parametrize("order,lapse,use_ml,OL,t", [(4, "Harmonic", True, 0.3, 0.2), (4, "Harmonic", False, 0.0, 0.0),
                                                   (6, "OnePlusLog", True, 0.0, 0.05), (8, "Static", True, 0.7, 1.0)])
def test_bardeen_parity(oracle_mod, gpu, order, lapse, use_ml, OL, t):
    _say(f"case bardeen order {order} lapse {lapse} ml {use_ml} OL {OL} t {t}")
    o, g, names = make_pair(oracle_mod, gpu, order, lapse)
    want_f, want_v, _ = oracle_bardeen(o, names, order, t, use_ml, OL)
    g.stepInit()
    got_v = g.bardeen(t, use_ml, OL)
    compare(g, want_f, want_v, got_v, 1e-10, 1e-9)
    assert np.abs(want_f["Phi"]).max() > 1e-6 and np.abs(want_f["D12"]).max() > 1e-6


def test_bardeen_shift_build(oracle_mod, gpu):
    """With the shift: h_0i, F, G_i and the shift terms of d^2 gamma~ / dt^2."""
    _say("case bardeen shift DampedWave order 4")
    o, g, names = make_pair(oracle_mod, gpu, 4, "DampedWave", shift="DampedWave")
    want_f, want_v, _ = oracle_bardeen(o, names, 4, 0.1, True, 0.0)
    g.stepInit()
    got_v = g.bardeen(0.1, True, 0.0)
    compare(g, want_f, want_v, got_v, 1e-10, 1e-9)
    assert np.abs(want_f["F"]).max() > 1e-8 and np.abs(want_f["G1"]).max() > 1e-8


def test_bardeen_after_fused_steps(oracle_mod, gpu):
    """After fused steps the evaluated state is stepInit's _a of the new _p
    (materialised by the output pass), as in the reference's output step."""
    _say("case bardeen after 2 fused steps")
    o, g, names = make_pair(oracle_mod, gpu, 4, "Harmonic")
    o.run_steps(2)
    g.step(2)
    want_f, want_v, _ = oracle_bardeen(o, names, 4, 0.3, True, 0.0)
    got_v = g.bardeen(0.3, True, 0.0)
    compare(g, want_f, want_v, got_v, 1e-10, 1e-9)
    # the evolution continues unchanged after the diagnostic
    o2, g2, _ = make_pair(oracle_mod, gpu, 4, "Harmonic")
    o2.run_steps(3)
    g.step(1)
    for f in FIELDS:
        assert rel_l2(g.download(f + "_p"), o2.field(f + "_p")) <= 1e-12, f


def test_bardeen_fp32_split(oracle_mod, gpu):
    _say("case bardeen fp32 split")
    o, g, names = make_pair(oracle_mod, gpu, 4, "Harmonic", precision="fp32_split")
    want_f, want_v, _ = oracle_bardeen(o, names, 4, 0.0, False, 0.0)
    g.stepInit()
    g.bardeen(0.0, False, 0.0)
    for name in ("Phi", "Psi", "A", "B", "D11", "D23", "C1"):
        assert rel_l2(g.download("Bardeen_" + name), want_f[name]) <= 1e-4, name


def test_bardeen_refuses_reference_metric(gpu):
    g = gpu.BSSN(gpu.GpuParams(n=16, order=4, matter="static"))
    g.frw_set(0.0, -3.0)
    g.stepInit()
    with pytest.raises(gpu.CosmoGpuError) as ei:
        g.bardeen(0.0)
    assert ei.value.code == -3
    with pytest.raises(gpu.CosmoGpuError):
        g.download("Bardeen_Phi")
