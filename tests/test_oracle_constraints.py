"""Pins the oracle's metric normalisation and constraint diagnostics
(oracle/oracle.c: or_bssn_normalize, point_constraints, or_bssn_constraints),
which restate BSSN::set_DIFFgamma_Aij_norm (bssn.cc:61-119) and
BSSN::setConstraintCalcs (bssn.cc:965-1219, bssn_macros.h:487-585).  The
reference's BSSN translation unit cannot be built here (it includes
<fftw3.h>), so these are checked against what the mathematics fixes:

  * stepInit's normalisation: det(delta + DIFFgamma) = 1 and
    gamma~^{ij} A_ij = 0 afterwards, and the closed form
    gamma^ = gamma / det^(1/3), A^ = (A - gamma^ tr(A)/3) / det^(1/3);
  * the Christoffel constraint G is zero when Gamma~^i = -d_j gamma~^{ij}
    (exact derivatives, unit determinant) and equals the injected violation
    otherwise;
  * the trace constraint A equals an injected trace 3c (A_ij += c gamma~_ij);
  * the unit-determinant constraint S equals |det(delta + DIFFgamma) - 1|;
  * the statistics (mean, stdev with N-1, max, sum of scales, scaled mean,
    stdev and max, L2 = sqrt(sum C^2 dx^3)) equal numpy's over the per-point
    values, and scaled = value / scale at every point.

The scale definitions (hamiltonianConstraintScale bssn.cc:1086-1100,
BSSN_MI_SCALE bssn_macros.h:510-527, BSSN_GI_SCALE bssn_macros.h:534-536)
are definitions without a mathematical counterpart; the S and A scales are
checked only on diagonal special cases, where they reduce to the sums of
absolute values of the terms they are built from."""
import math

import numpy as np
import pytest

from common import FIELDS, smooth_random_state

torch = pytest.importorskip("torch")
from test_oracle_adm import SYM, Data, oracle_point  # noqa: E402

pytestmark = []


def sym_np(v):
    return np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]])


def random_state_sim(po, shape, seed, amp_g=0.05, amp=1e-2):
    """An oracle Sim holding a smooth random state whose metric is far from
    unit determinant and whose A_ij carries a trace."""
    st = smooth_random_state(shape, seed=seed, amp=amp)
    rng = np.random.default_rng(seed + 100)
    for f in FIELDS[:6]:
        st[f] = st[f] * (amp_g / amp) + (0.02 * rng.standard_normal() if f in ("DIFFgamma11", "DIFFgamma33") else 0.0)
    s = po.Sim(po.make_params(shape, order=4))
    for f in FIELDS:
        s.field(f + "_p")[:] = st[f]
    s.field("DIFFD_a")[:] = st["DIFFD"]
    s.frw_set(0.0, -3.0)
    s.frw_add_fluid(3.0 / (8.0 * math.pi), 0.0)
    return s, st


def test_step_init_normalises_metric_and_trace(oracle_mod):
    shape = (10, 12, 8)
    s, st = random_state_sim(oracle_mod, shape, seed=7)
    s.stepInit()
    g = np.stack([s.field(f + "_a") for f in FIELDS[:6]], axis=-1).reshape(-1, 6)
    A = np.stack([s.field(f + "_a") for f in FIELDS[7:13]], axis=-1).reshape(-1, 6)
    g0 = np.stack([st[f] for f in FIELDS[:6]], axis=-1).reshape(-1, 6)
    A0 = np.stack([st[f] for f in FIELDS[7:13]], axis=-1).reshape(-1, 6)
    worst_det, worst_tr, worst_g, worst_A = 0.0, 0.0, 0.0, 0.0
    for x in range(0, g.shape[0], 7):
        m = np.eye(3) + sym_np(g[x])
        worst_det = max(worst_det, abs(np.linalg.det(m) - 1.0))
        a = sym_np(A[x])
        worst_tr = max(worst_tr, abs(np.trace(np.linalg.inv(m) @ a)) / np.abs(a).max())
        # closed form from the unnormalised input
        m0 = np.eye(3) + sym_np(g0[x])
        c = np.cbrt(np.linalg.det(m0))
        gh = m0 / c
        a0 = sym_np(A0[x])
        ah = (a0 - gh * np.trace(np.linalg.inv(gh) @ a0) / 3.0) / c
        worst_g = max(worst_g, np.abs(gh - m).max())
        worst_A = max(worst_A, np.abs(ah - a).max() / np.abs(ah).max())
    assert worst_det < 1e-14, worst_det
    assert worst_tr < 1e-14, worst_tr
    assert worst_g < 1e-15, worst_g
    assert worst_A < 1e-13, worst_A
    # stepInit leaves the other fields' _a equal to _p, and _f zero
    for f in FIELDS[6:7] + FIELDS[13:]:
        assert np.array_equal(s.field(f + "_a"), s.field(f + "_p")), f
    for f in FIELDS:
        assert not np.any(s.field(f + "_f")), f
    # a wrong-sign trace removal or a missing 1/det^(1/3) fails the closed form:
    # the input itself is far from it
    assert np.abs(A0 - A).max() > 1e3 * worst_A * np.abs(A).max()


@pytest.mark.parametrize("seed", [1, 2])
def test_point_gamma_constraint(oracle_mod, seed):
    x = torch.tensor(np.random.default_rng(300 + seed).uniform(0, 1, 3))
    params = oracle_mod.make_params(16, order=4)
    d0, d1 = Data(seed), Data(seed, dgam_amp=0.02)
    c0 = oracle_mod.point_constraints(oracle_point(oracle_mod, params, d0, x))
    c1 = oracle_mod.point_constraints(oracle_point(oracle_mod, params, d1, x))
    # consistent Gamma~^i = -d_j gamma~^{ij}: zero to round-off
    assert c0[2, 0] < 1e-13, c0[2, 0]
    # injected violation delta Gamma^i: G = |delta Gamma|
    inj = (d1.Gam(x) - d0.Gam(x)).detach().numpy()
    want = float(np.sqrt(np.sum(inj * inj)))
    assert abs(c1[2, 0] - want) <= 1e-12 * want, (c1[2, 0], want)
    for c in (c0, c1):
        for q in range(5):
            assert c[q, 2] == c[q, 0] / c[q, 1]


def test_point_trace_constraint(oracle_mod):
    x = torch.tensor([0.13, 0.58, 0.91])
    params = oracle_mod.make_params(16, order=4)
    d = Data(3)
    pt = oracle_point(oracle_mod, params, d, x)
    assert abs(oracle_mod.point_constraints(pt)[3, 0]) < 1e-15  # A~ is trace-free by construction
    gt = d.gt(x).detach().numpy()
    cval = 0.0371
    for q, (a, b) in enumerate(SYM):
        pt.v[7 + q] += cval * gt[a, b]
    oracle_mod.point_compute(params, pt)
    got = oracle_mod.point_constraints(pt)[3, 0]
    assert abs(got - 3 * cval) <= 1e-14, got
    # flat metric, diagonal special case of the scale: |A11| + |A22| + |A33|
    pt2 = oracle_mod.Point()
    for q, v in enumerate([0.3, 0.7, -0.2, -0.5, 0.4, 0.9]):
        pt2.v[7 + q] = v
    oracle_mod.point_compute(params, pt2)
    c = oracle_mod.point_constraints(pt2)
    assert c[3, 0] == 0.3 - 0.5 + 0.9
    assert c[3, 1] == 0.3 + 0.5 + 0.9


def test_point_unit_det_constraint(oracle_mod):
    params = oracle_mod.make_params(16, order=4)
    rng = np.random.default_rng(11)
    for _ in range(5):
        gv = 0.1 * rng.standard_normal(6)
        pt = oracle_mod.Point()
        for q in range(6):
            pt.v[q] = gv[q]
        oracle_mod.point_compute(params, pt)
        got = oracle_mod.point_constraints(pt)[4, 0]
        want = abs(np.linalg.det(np.eye(3) + sym_np(gv)) - 1.0)
        assert abs(got - want) <= 1e-15, (got, want)
    # diagonal special case of the scale: the absolute values of the terms
    # of det(1 + diag(e)) - 1 = sum e_i + sum e_i e_j + e1 e2 e3
    e = [0.1, -0.2, 0.05]
    pt = oracle_mod.Point()
    pt.v[0], pt.v[3], pt.v[5] = e
    oracle_mod.point_compute(params, pt)
    c = oracle_mod.point_constraints(pt)
    terms = [e[0], e[1], e[2], e[0] * e[1], e[0] * e[2], e[1] * e[2], e[0] * e[1] * e[2]]
    assert abs(c[4, 0] - abs(sum(terms))) < 1e-16
    assert abs(c[4, 1] - sum(abs(t) for t in terms)) < 1e-16


def test_constraint_statistics_match_numpy(oracle_mod):
    """setConstraintCalcs' statistics over the per-point values, and the H
    and M_1 values against the reference's 1D strip output reassembled."""
    shape = (12, 10, 14)
    s, _ = random_state_sim(oracle_mod, shape, seed=3, amp_g=1e-2)
    s.stepInit()
    s.clearSrc()
    L = oracle_mod.lib()
    L.or_static_add_src(s.h)
    L.or_lambda_add_src(s.h)
    got = s.constraints()
    pts = s.constraint_points()
    dx = s.p.dx
    for q, name in enumerate("HMGAS"):
        val, sca, scd = (pts[..., q, k].ravel() for k in range(3))
        np.testing.assert_array_equal(scd, val / sca)
        want = [val.mean(), val.std(ddof=1), np.abs(val).max(), sca.sum(), scd.mean(), scd.std(ddof=1),
                np.abs(scd).max()]
        if q < 2:
            want.append(math.sqrt(np.sum(val * val) * dx ** 3))
        for k, w in enumerate(want):
            assert abs(got[name][k] - w) <= 1e-13 * max(abs(w), np.abs(val).max() * 1e-3), (name, k, got[name][k], w)
        assert np.abs(val).max() > 0, name
    # reassembled strips (set1DConstraintOutput along i at every (j, k)):
    # H and the M_1 component, whose square is part of |M|^2
    nx, ny, nz = shape
    for j in range(0, ny, 3):
        for k in range(0, nz, 4):
            H, M1 = s.constraint_strip(1, j, k)
            np.testing.assert_array_equal(H, pts[:, j, k, 0, 0])
            assert np.all(np.abs(M1) <= pts[:, j, k, 1, 0] * (1 + 1e-15))
