"""Output-step readers on the GPU against the oracle: field strips
(io_dump_strip, IO/io.cc:719-771), the 1D constraint strips
(BSSN::set1DConstraintOutput, bssn.cc:903-938), the conformal statistics of
io_bssn_dump_statistics (IO/io.cc:8-14, 341-403) and the sources the output
step sees (DIFFr_a / DIFFS_a after the constraint pass)."""
import sys

import numpy as np
import pytest

from common import FIELDS, setup_pair, smooth_random_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _output_point(po, o, matter):
    """The oracle at the output step: stepInit, clearSrc, addBSSNSrc."""
    L = po.lib()
    L.or_bssn_step_init(o.h)
    if matter == "dust":
        L.or_dust_step_init(o.h)
    L.or_bssn_clear_src(o.h)
    (L.or_dust_add_src if matter == "dust" else L.or_static_add_src)(o.h)
    L.or_lambda_add_src(o.h)


@pytest.mark.parametrize("prec", ["fp64", "fp32_split"])
def test_field_strips_are_slices(gpu, prec):
    shape = (12, 16, 20)
    _say(f"case strips {prec} shape={shape}")
    g = gpu.BSSN(gpu.GpuParams(n=shape, order=4, precision=prec))
    st = smooth_random_state(shape, seed=11)
    for f in FIELDS:
        g.upload(f + "_p", st[f])
    g.upload("DIFFD_a", st["DIFFD"])
    for name in ("DIFFphi_p", "A12_p", "DIFFK_a", "DIFFD_a", "Gamma3_p"):
        full = g.download(name)
        for n1, n2 in ((0, 0), (5, 7), (-1, 23), (30, -45)):
            a = full[:, n1 % shape[1], n2 % shape[2]]
            b = full[n1 % shape[0], :, n2 % shape[2]]
            c = full[n1 % shape[0], n2 % shape[1], :]
            assert np.array_equal(g.strip(name, 1, n1, n2), a), (name, 1, n1, n2)
            assert np.array_equal(g.strip(name, 2, n1, n2), b), (name, 2, n1, n2)
            assert np.array_equal(g.strip(name, 3, n1, n2), c), (name, 3, n1, n2)


@pytest.mark.parametrize("matter,shape", [("static", (16, 18, 20)), ("dust", (20, 16, 18))])
def test_constraint_strips_and_sources_match_oracle(oracle_mod, gpu, matter, shape):
    _say(f"case constraint strips {matter} shape={shape}")
    st = smooth_random_state(shape, seed=21, amp=3e-3, with_alpha=False)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4, "matter": matter, "lam": 0.01}, state=st)
    if matter == "dust":
        D = 0.05 + 0.1 * st["DIFFD"]
        for name, arr in (("dust_D_p", D), ("dust_S1_p", 0.3 * st["DIFFK"]), ("dust_S3_p", 0.2 * st["A12"])):
            o.field(name)[:] = arr
            g.upload(name, arr)
    o.run_steps(1)
    g.step(1)
    _output_point(oracle_mod, o, matter)
    g.setConstraintCalcs()
    # the sources of this state
    r_want = o.field("DIFFr_a")
    np.testing.assert_allclose(g.download("DIFFr_a"), r_want, rtol=1e-11, atol=1e-12 * np.abs(r_want).max())
    S_want = o.field("DIFFS_a")
    np.testing.assert_allclose(g.download("DIFFS_a"), S_want, rtol=1e-11, atol=1e-12 * max(np.abs(S_want).max(), 1e-3))
    for axis, n1, n2 in ((1, 0, 0), (1, 5, 9), (2, 3, 4), (3, -2, 40)):
        Hg, Mg = g.set1DConstraintOutput(axis, n1, n2)
        Ho, Mo = o.constraint_strip(axis, n1, n2)
        assert Hg.size == shape[axis - 1]
        hs, ms = np.abs(Ho).max(), np.abs(Mo).max()
        assert hs > 0 and ms > 0
        np.testing.assert_allclose(Hg, Ho, rtol=0, atol=1e-10 * hs, err_msg=f"H axis {axis}")
        np.testing.assert_allclose(Mg, Mo, rtol=0, atol=1e-10 * ms, err_msg=f"M axis {axis}")


@pytest.mark.parametrize("prec", ["fp64", "fp32_split"])
def test_field_stats_match_oracle(oracle_mod, gpu, prec):
    shape = (16, 16, 20)
    _say(f"case field stats {prec} shape={shape}")
    st = smooth_random_state(shape, seed=31, amp=3e-3, with_alpha=False)
    o, g = setup_pair(oracle_mod, gpu, {"n": shape, "order": 4, "precision": prec}, state=st)
    o.run_steps(1)
    g.step(1)
    _output_point(oracle_mod, o, "static")
    oracle_mod.lib().or_bssn_prep_output(o.h)
    g.setConstraintCalcs()
    tol = 1e-12 if prec == "fp64" else 1e-5
    for name in ("DIFFphi_a", "DIFFK_a", "DIFFr_a", "ricci_a"):
        got, want = g.field_stats(name), o.conformal_stats(name)
        scale = np.abs(o.field(name)).max()
        assert abs(got[0] - want[0]) <= tol * scale, (name, got, want)
        assert abs(got[1] - want[1]) <= tol * max(want[1], 1e-3 * scale), (name, got, want)
        assert abs(got[2] - want[2]) <= 1e-12 * want[2], (name, got, want)


def test_output_readers_reject_bad_requests(gpu):
    """Error paths of the readers: a strip must hold the grid's extent along
    its axis, the axis is 1..3, names must exist, and the sources of a
    static-matter context exist only after the constraint pass."""
    import ctypes as C
    shape = (16, 16, 20)
    _say("case output reader errors")
    g = gpu.BSSN(gpu.GpuParams(n=shape, order=4))
    L = g.L
    buf = np.zeros(64)
    ptr = buf.ctypes.data_as(C.POINTER(C.c_double))
    assert L.cg_field_strip(g.h, b"DIFFphi_p", 1, 0, 0, ptr, 15) == -1          # CG_ERR_ARG: NX is 16
    assert L.cg_field_strip(g.h, b"DIFFphi_p", 4, 0, 0, ptr, 16) == -1          # no axis 4
    assert L.cg_field_strip(g.h, b"nosuch_p", 1, 0, 0, ptr, 16) == -1
    assert L.cg_field_strip(g.h, b"DIFFr_a", 1, 0, 0, ptr, 16) == -3            # CG_ERR_STATE before the pass
    assert L.cg_constraint_strip(g.h, 3, 0, 0, ptr, ptr, 16) == -1              # NZ is 20
    out3 = (C.c_double * 3)()
    assert L.cg_field_stats(g.h, b"nosuch_a", out3) == -1
    assert L.cg_field_stats(g.h, b"ricci_a", out3) == -3
    g.setConstraintCalcs()
    assert L.cg_field_strip(g.h, b"DIFFr_a", 3, 0, 0, ptr, 20) == 0
    assert L.cg_field_stats(g.h, b"ricci_a", out3) == 0


@pytest.mark.parametrize("prec,shape", [("fp64", (16, 16, 20)), ("fp32_split", (20, 12, 16))])
def test_power_spectrum_matches_oracle(oracle_mod, gpu, prec, shape):
    _say(f"case power spectrum {prec} shape={shape}")
    st = smooth_random_state(shape, seed=41, kmax=4)
    g = gpu.BSSN(gpu.GpuParams(n=shape, order=4, precision=prec))
    for f in FIELDS:
        g.upload(f + "_p", st[f])
    for name in ("DIFFphi_p", "DIFFK_p", "A13_p"):
        want = oracle_mod.power_dump(g.download(name))
        got = g.power_spectrum(name)
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12 * want.max(), err_msg=name)
