"""Output-step quantities of the oracle: the 1D constraint strips of
BSSN::set1DConstraintOutput (bssn.cc:903-938).

The strips are pinned by composition.  The reference passes the running index
as i for every axis (bssn.cc:921, 931), so on a cubic grid the strips along
axes 1, 2 and 3 at (n1, n2) all run over the points (x, n1, n2) and hold M_1,
M_2 and M_3 there.  The strips of every (n1, n2) therefore reassemble H and
|M| on the whole grid, whose mean, max and L2 must equal setConstraintCalcs'
(bssn.cc:965-1071; its point algebra is pinned by the ADM identity tests).
A strip that evaluated another point, another component or another register
fails this.  The conformal standard deviation of the statistics file is
pinned bitwise against the reference in test_oracle_ref.py."""
import numpy as np
import pytest

from common import FIELDS, smooth_random_state


def output_state(po, shape, seed, matter="static"):
    """An oracle sim at the reference's output point: stepInit, clearSrc,
    addBSSNSrc (initStaticStep, sims/static.cc:74-82)."""
    s = po.Sim(po.make_params(shape, order=4, matter=matter))
    st = smooth_random_state(shape, seed=seed, amp=3e-3, with_alpha=False)
    for f in FIELDS:
        s.field(f + "_p")[:] = st[f]
    s.field("DIFFD_a")[:] = 0.05 * st["DIFFD"]
    s.frw_set(0.0, -3.0)
    s.frw_add_fluid(3.0 / (8.0 * np.pi), 0.0)
    L = po.lib()
    L.or_bssn_step_init(s.h)
    L.or_bssn_clear_src(s.h)
    L.or_static_add_src(s.h)
    L.or_lambda_add_src(s.h)
    return s


def test_constraint_strips_reassemble_the_constraint_statistics(oracle_mod):
    n = 10
    s = output_state(oracle_mod, (n, n, n), seed=3)
    H = np.zeros((n, n, n))
    M = np.zeros((3, n, n, n))
    for n1 in range(n):
        for n2 in range(n):
            for axis in (1, 2, 3):
                h, m = s.constraint_strip(axis, n1, n2)
                if axis == 1:
                    H[:, n1, n2] = h
                else:
                    np.testing.assert_array_equal(h, H[:, n1, n2])  # the same points for every axis
                M[axis - 1, :, n1, n2] = m
    want = s.constraints()
    dx3 = s.p.dx ** 3
    Mn = np.sqrt((M ** 2).sum(axis=0))
    for c, v in (("H", H), ("M", Mn)):
        w = want[c]
        assert abs(np.mean(v) - w[0]) <= 1e-12 * np.abs(v).max(), c
        assert np.abs(v).max() == pytest.approx(w[2], rel=1e-14), c
        assert np.sqrt(np.sum(v * v * dx3)) == pytest.approx(w[7], rel=1e-12), c
    assert np.abs(H).max() > 0 and np.abs(Mn).max() > 0


def test_constraint_strip_runs_along_i_on_every_axis(oracle_mod):
    """On a non-cubic grid the axis-2 strip holds NY values at (x mod NX, n1, n2)."""
    shape = (8, 10, 12)
    s = output_state(oracle_mod, shape, seed=4)
    h1, _ = s.constraint_strip(1, 3, 5)
    h2, _ = s.constraint_strip(2, 3, 5)
    h3, _ = s.constraint_strip(3, 3 + 10, 5 - 12)   # offsets wrap periodically
    assert h1.size == 8 and h2.size == 10 and h3.size == 12
    np.testing.assert_array_equal(h2[:8], h1)
    np.testing.assert_array_equal(h2[8:], h1[:2])
    np.testing.assert_array_equal(h3, np.concatenate([h1, h1[:4]]))


def test_power_dump_weights_satisfy_parseval(oracle_mod):
    """The bins' counts times their averages sum to sum_k |F_k|^2 over the full
    spectrum, which Parseval fixes at N^3 sum_x f^2: a wrong weight for the
    pz = 0 / NZ/2 planes or for the conjugate modes breaks it."""
    shape = (12, 10, 14)
    f = np.random.default_rng(5).standard_normal(shape)
    P = oracle_mod.power_dump(f)
    nx, ny, nz = shape
    px = np.where(np.arange(nx) <= nx // 2, np.arange(nx), np.arange(nx) - nx)
    py = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny)
    PX, PY, PZ = np.meshgrid(px, py, np.arange(nz // 2 + 1), indexing="ij")
    bins = np.sqrt((PX * PX + PY * PY + PZ * PZ).astype(float)).astype(int)
    w = np.where((PZ >= 1) & (PZ < nz // 2), 2, 1)
    cnt = np.bincount(bins.ravel(), weights=w.ravel(), minlength=P.size)
    assert cnt.sum() == f.size                      # every mode of the full spectrum once
    total = float(np.sum(cnt * P))
    assert total == pytest.approx(f.size * float(np.sum(f * f)), rel=1e-12)


def test_power_dump_of_a_single_mode(oracle_mod):
    """cos(2 pi (3x + 4y)/N) has all its power in the bin |p| = 5."""
    n = 16
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = np.cos(2 * np.pi * (3 * X + 4 * Y) / n)
    P = oracle_mod.power_dump(f)
    assert np.argmax(P) == 5
    others = np.delete(P, 5)
    assert np.max(others) <= 1e-20 * P[5]
