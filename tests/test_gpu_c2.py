"""The 128^3 FP64 static-dust configuration at its full size against the
oracle: 200 RK4 steps of order-4 BSSN + static dust from the seeded
constraint-solved ICs (P(k) ~ k^-2, rms 1e-3, seed 9).  The oracle's run is
held in tests/golden/c2_digest.json and c2_sample.npz
(tests/golden/make_digest.py): the H and M statistics at the output point of
every 10th step; per field the L2 norm and largest magnitude of _p after 200
steps, and _p on the strided 32^3 sub-lattice.

Bars (the FP64 drop-in's 1e-12):
  * every field: relative L2 of the difference on the 32^3 sample, and the
    full field's L2 norm and largest magnitude, <= 1e-12 (fields the oracle
    holds at zero: exactly zero);
  * constraint statistics: the sum of scales to 1e-12 relative; the value
    statistics (mean, stdev, max, L2) to 1e-12 of max(|value|, s), s the
    mean scale (H[3]/N, M[3]/N), i.e. of the magnitude of the terms whose
    difference the constraint is (H = R + 2/3 K^2 - A.A - 16 pi rho is a
    residual ~1e-5 of terms ~1 at the first steps); the scaled statistics to
    1e-12 of max(|value|, 1) for the same reason;
  * the FRW background bit for bit.
The GPU path runs its default (accumulator-free) RK4 form; the oracle runs
the reference's."""
import sys

import numpy as np
import pytest

import digests
from common import FIELDS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_constraints(got, rec, npts, worst):
    for c in ("H", "M"):
        gaps, want = digests.constraint_gap(got[c], rec[c])
        s = abs(want[3]) / npts
        for q, w in enumerate(want):
            worst[c][q] = max(worst[c][q], gaps[q])
            if np.isnan(w):
                assert np.isnan(got[c][q]), (rec["step"], c, q, got[c][q])
                continue
            floor = {3: 0.0, 4: 1.0, 5: 1.0, 6: 1.0}.get(q, s)
            assert abs(got[c][q] - w) <= 1e-12 * max(abs(w), floor), (rec["step"], c, q, got[c][q], w)


def test_c2_full_size_parity(gpu):
    from cosmograph_amd.ics import static_dust_ics
    d, sample = digests.load("c2")
    cfg = d["config"]
    n = cfg["n"]
    _say(f"case c2 full size n={n} order={cfg['order']} steps={cfg['steps']}")
    g = digests.gpu_context(gpu, d, "fp64", static_dust_ics)
    worst = {"H": [0.0] * 8, "M": [0.0] * 8}
    for rec in d["constraints"]:
        if rec["step"]:
            g.step(cfg["every"])
        check_constraints(g.setConstraintCalcs(), rec, float(n) ** 3, worst)
    _say(f"c2 worst relative constraint gaps: H {['%.1e' % x for x in worst['H']]} "
         f"M {['%.1e' % x for x in worst['M']]}")
    assert list(g.frw_get()) == [digests.fhex(x) for x in d["frw"]]
    errs = digests.sample_errors(g, d, sample)
    _say("c2 field gaps (sample rel L2): " + " ".join(f"{f} {errs[f]['sample']:.1e}" for f in FIELDS))
    for f in FIELDS:
        for k in ("sample", "l2", "max"):
            assert errs[f][k] <= 1e-12, (f, k, errs[f][k])
