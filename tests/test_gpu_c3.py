"""The 256^3 split-precision configuration at its full size against the
FP64 oracle: 200 RK4 steps of order-6 BSSN + static dust from the seeded
constraint-solved ICs (P(k) ~ k^-3 on 1 <= |k| <= N/4, rms 1e-2, seed 9,
dt = 0.1 dx).  The oracle's run is held in tests/golden/c3_digest.json and
c3_sample.npz (tests/golden/make_digest.py, about 2.5 hours on 8 cores).

FP64 drop-in: the bars of the 128^3 FP64 test (1e-12 on every field's
sample and norms; constraint statistics to 1e-12 of the terms they are the
difference of).

The two 4-byte modes, FP32 split (FP32 storage and stencils of A_ij, K,
Gamma^i, alpha, D; FP64 conformal metric) and the fixed-point metric (the
same with the conformal metric in 32-bit words), both with FP64 RK4
accumulation, FRW and reductions:
  * every field within 1e-5 relative L2 of the oracle on the 32^3 sample;
  * the Hamiltonian constraint no worse than the oracle's, to a stated
    tolerance: H L2 and H L-infinity at every output step at most
    (1 + 1e-4) times the oracle's plus four FP32 ulps of the mean scale of
    H (the size of the terms it is the difference of).  Rounding the stored
    fields to 4 bytes adds a violation of its own, about 1e-5 of H here once
    it has grown (split mode: +1.0e-5 in L2, +1.9e-5 in L-infinity after 203
    steps); the relative tolerance allows 5x that, the floor covers the
    first output steps, where H is the constraint-solved initial residual,
    and a drift of the evolution (growth of H) would exceed both."""
import sys

import numpy as np
import pytest

import digests
from common import FIELDS
from test_gpu_c2 import check_constraints

pytestmark = pytest.mark.gpu

H_TOL = 1e-4
# Absolute floor of the H bar, in units of the mean scale of H per point (the
# magnitude of the terms H is the difference of, ~1.06 here): four FP32 ulps.
# It matters only while H is tiny: the initial data are constraint-solved (H
# L-infinity 5e-5 at step 0), and rounding the stored state to 4 bytes
# (the conformal metric's 32-bit words included) moves that residual's
# maximum by ~1e-7 of the scale, i.e. by up to 2e-3 of its own value.
H_FLOOR = 4 * 2.0 ** -23


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


@pytest.fixture(scope="module")
def digest():
    try:
        return digests.load("c3")
    except FileNotFoundError:
        pytest.skip("tests/golden/c3_digest.json not generated (tests/golden/make_digest.py c3)")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_c3_fp64_parity(gpu, digest):
    from cosmograph_amd.ics import static_dust_ics
    d, sample = digest
    cfg = d["config"]
    n = cfg["n"]
    _say(f"case c3 fp64 n={n} order={cfg['order']} steps={cfg['steps']}")
    g = digests.gpu_context(gpu, d, "fp64", static_dust_ics)
    worst = {"H": [0.0] * 8, "M": [0.0] * 8}
    for rec in d["constraints"]:
        if rec["step"]:
            g.step(cfg["every"])
        check_constraints(g.setConstraintCalcs(), rec, float(n) ** 3, worst)
    _say(f"c3 fp64 worst relative constraint gaps: H {['%.1e' % x for x in worst['H']]} "
         f"M {['%.1e' % x for x in worst['M']]}")
    assert list(g.frw_get()) == [digests.fhex(x) for x in d["frw"]]
    errs = digests.sample_errors(g, d, sample)
    _say("c3 fp64 field gaps (sample rel L2): " + " ".join(f"{f} {errs[f]['sample']:.1e}" for f in FIELDS))
    for f in FIELDS:
        for k in ("sample", "l2", "max"):
            assert errs[f][k] <= 1e-12, (f, k, errs[f][k])


@pytest.mark.parametrize("prec", ["fp32_split", "fp32_fix"])
def test_c3_fp32_split_gate(gpu, digest, prec):
    """The same gate for both 4-byte-perturbation modes: the split mode (FP64
    conformal metric) and the fixed-point metric mode (the metric in 32-bit
    words, 4 bytes per value like every other field)."""
    from cosmograph_amd.ics import static_dust_ics
    d, sample = digest
    cfg = d["config"]
    _say(f"case c3 {prec} n={cfg['n']} order={cfg['order']} steps={cfg['steps']}")
    g = digests.gpu_context(gpu, d, prec, static_dust_ics)
    excess = {"L2": 0.0, "Linf": 0.0}
    bad = []
    npts = float(cfg["n"]) ** 3
    for rec in d["constraints"]:
        if rec["step"]:
            g.step(cfg["every"])
        got = g.setConstraintCalcs()["H"]
        want = [digests.fhex(x) for x in rec["H"]]
        floor = H_FLOOR * want[3] / npts  # want[3]: the sum of the H scale over the grid
        for key, q in (("L2", 7), ("Linf", 2)):
            r = got[q] / want[q] - 1.0
            excess[key] = max(excess[key], r)
            if got[q] > want[q] * (1.0 + H_TOL) + floor:
                bad.append((rec["step"], key, r, got[q] - want[q], floor))
    assert not bad, bad
    errs = digests.sample_errors(g, d, sample)
    _say(f"c3 {prec} largest relative H excess over the oracle: L2 {excess['L2']:.2e}, Linf {excess['Linf']:.2e}")
    _say(f"c3 {prec} field errors (sample rel L2): " + " ".join(f"{f} {errs[f]['sample']:.1e}" for f in FIELDS))
    for f in FIELDS:
        assert errs[f]["sample"] <= 1e-5, (f, errs[f]["sample"])
