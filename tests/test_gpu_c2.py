"""The 128^3 FP64 static-dust configuration at its full size against the
oracle: 200 RK4 steps of order-4 BSSN + static dust from the seeded
constraint-solved ICs (P(k) ~ k^-2, rms 1e-3, seed 9).  The oracle's run is
held in tests/golden/c2_digest.json (tests/golden/make_c2_digest.py): the H
and M statistics at the output point of every 10th step, and per field the
sum, L2 norm, largest magnitude and 64 probe values of _p after 200 steps.

Bars: field L2 norms and largest magnitudes <= 1e-12 relative, sums within
1e-12 of the L2 bound sqrt(N^3)*L2, probe values within 1e-11 of the field's
largest magnitude, every constraint statistic within 1e-9 (relative, floored
at 1e-3 of the largest |H| or |M|; the constraints differentiate the state
twice), the FRW background bit for bit.  The GPU path runs its default
(accumulator-free) RK4 form; the oracle runs the reference's."""
import json
import os
import sys

import numpy as np
import pytest

from common import FIELDS

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _f(h):
    return float.fromhex(h)


def test_c2_full_size_parity(gpu):
    from cosmograph_amd.ics import static_dust_ics
    D = json.load(open(os.path.join(HERE, "golden", "c2_digest.json")))
    cfg = D["config"]
    n = cfg["n"]
    _say(f"case c2 full size n={n} order={cfg['order']} steps={cfg['steps']}")
    fields, frw = static_dust_ics(n, cfg["spectral_index"], cfg["rms"], cfg["seed"])
    g = gpu.BSSN(gpu.GpuParams(n=n, order=cfg["order"], precision="fp64"))
    for name, arr in fields.items():
        g.upload(name, arr)
    g.frw_set(frw["phi"], frw["K"])
    for rho, w in frw["fluids"]:
        g.frw_add_fluid(rho, w)
    for rec in D["constraints"]:
        if rec["step"]:
            g.step(cfg["every"])
        got = g.setConstraintCalcs()
        for c in ("H", "M"):
            want = [_f(x) for x in rec[c]]
            for q in range(8):
                w = want[q]
                if np.isnan(w):  # the scaled statistics of a constraint that vanishes (M of the ICs: 0/0)
                    assert np.isnan(got[c][q]), (rec["step"], c, q, got[c][q])
                    continue
                assert abs(got[c][q] - w) <= 1e-9 * max(abs(w), 1e-3 * abs(want[2])), (rec["step"], c, q, got[c][q], w)
    assert list(g.frw_get()) == [_f(x) for x in D["frw"]]
    pts = D["probe_points"]
    npts = float(n) ** 3
    for f in FIELDS:
        a = g.download(f + "_p")
        d = D["fields"][f]
        l2, mx = _f(d["l2"]), _f(d["max_abs"])
        if l2 == 0.0:
            assert not np.any(a), f
            continue
        assert abs(np.sqrt(np.sum(a * a)) - l2) <= 1e-12 * l2, f
        assert abs(np.abs(a).max() - mx) <= 1e-12 * mx, f
        assert abs(a.sum() - _f(d["sum"])) <= 1e-12 * np.sqrt(npts) * l2, f
        probe = np.array([a[i, j, k] for i, j, k in pts])
        want = np.array([_f(x) for x in d["probes"]])
        assert np.max(np.abs(probe - want)) <= 1e-11 * mx, (f, np.max(np.abs(probe - want)) / mx)
