"""Batched ensemble: many grids advanced by one stage-kernel launch per RK4
stage, replayed from a HIP graph.  Every realisation must evolve exactly as
it would alone: against the oracle (FP64 <= 1e-12 relative L2, FP32 split
<= 1e-5), against the single-grid path, and graph replays against eager
launches, with eager work between the replays."""
import sys

import numpy as np
import pytest

from common import FIELDS, rel_l2, setup_pair, smooth_random_state

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (18, 14, 20), (24, 24, 24), (16, 16, 16)]
FRW = {"phi": 0.0, "K": -3.0, "fluids": [(3.0 / (8.0 * np.pi), 0.0)]}


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _make_batch(cg, shapes, precision="fp64", order=4, lam=(0.0,), seeds=None, graph=True):
    params = [cg.GpuParams(n=s, order=order, precision=precision, lambda_=lam[r % len(lam)])
              for r, s in enumerate(shapes)]
    b = cg.Batch(params)
    b.set_graph(graph)
    states = []
    for r, (m, s) in enumerate(zip(b.members, shapes)):
        st = smooth_random_state(s, seed=(seeds[r] if seeds else 100 + r), with_alpha=False)
        for f in FIELDS:
            m.upload(f + "_p", st[f])
        m.upload("DIFFD_a", st["DIFFD"])
        m.frw_set(FRW["phi"], FRW["K"])
        for rho, w in FRW["fluids"]:
            m.frw_add_fluid(rho, w)
        states.append(st)
    return b, states


def _oracle(oracle_mod, gpu, shape, st, order=4, lam=0.0):
    o, _ = setup_pair(oracle_mod, gpu, {"n": shape, "order": order, "lam": lam}, state=st)
    return o


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32_split", 1e-5)])
def test_batch_matches_oracle(oracle_mod, gpu, precision, tol):
    _say(f"case batch vs oracle {precision} shapes={SHAPES}")
    lam = (0.0, 0.01)
    b, states = _make_batch(gpu, SHAPES, precision, lam=lam)
    b.step(2)
    b.step(1)   # a second captured length
    b.step(2)   # replay of the first graph
    for r, (shape, st) in enumerate(zip(SHAPES, states)):
        o = _oracle(oracle_mod, gpu, shape, st, lam=lam[r % 2])
        o.run_steps(5)
        m = b.members[r]
        for f in FIELDS:
            err = rel_l2(m.download(f + "_p"), o.field(f + "_p"))
            assert err <= tol, (r, f, err)
        assert np.allclose(m.frw_get(), o.frw_get(), rtol=0, atol=0), r
        ex, want = m.extra_field_data(), o.extra()
        for k in ("K_min", "K_avg", "avg_vol", "rho_avg"):
            assert abs(ex[k] - want[k]) <= (1e-12 if precision == "fp64" else 1e-5) * abs(want[k]), (r, k)


def test_batch_graph_replays_match_eager(gpu):
    """Graph replays, with eager work (downloads, a constraint evaluation)
    between them, against the same launches issued one by one."""
    _say("case batch graph vs eager")
    bg, _ = _make_batch(gpu, SHAPES, "fp32_split", graph=True)
    be, _ = _make_batch(gpu, SHAPES, "fp32_split", graph=False)
    for call in range(3):
        bg.step(2)
        be.step(2)
        bg.setConstraintCalcs()
        for r in range(len(SHAPES)):
            for f in FIELDS:
                a = bg.members[r].download(f + "_p")
                e = be.members[r].download(f + "_p")
                assert np.array_equal(a, e), (call, r, f)


def test_batch_matches_single_grid(gpu):
    _say("case batch vs single grid")
    b, states = _make_batch(gpu, SHAPES, "fp64")
    b.step(3)
    for r, (shape, st) in enumerate(zip(SHAPES, states)):
        g = gpu.BSSN(gpu.GpuParams(n=shape, order=4))
        for f in FIELDS:
            g.upload(f + "_p", st[f])
        g.upload("DIFFD_a", st["DIFFD"])
        g.frw_set(FRW["phi"], FRW["K"])
        for rho, w in FRW["fluids"]:
            g.frw_add_fluid(rho, w)
        g.step(3)
        for f in FIELDS:
            assert rel_l2(b.members[r].download(f + "_p"), g.download(f + "_p")) <= 1e-14, (r, f)


def test_batch_tile_orders_give_identical_results(gpu):
    """The strip tile orders (4, 5) in the batched kernel, with partial strips
    (2 and 3 tile rows along j), equal the default order bit for bit."""
    shapes = [(16, 16, 16), (20, 16, 24), (24, 24, 24)]
    _say(f"case batch tile orders shapes={shapes}")
    out = {}
    for order in (0, 4, 5):
        b, _ = _make_batch(gpu, shapes, "fp32_split")
        for m in b.members:
            m.set_tile_order(order)
        b.step(2)
        out[order] = [{f: m.download(f + "_p") for f in FIELDS} for m in b.members]
        del b
    for order in (4, 5):
        for r in range(len(shapes)):
            for f in FIELDS:
                assert np.array_equal(out[order][r][f], out[0][r][f]), (order, r, f)


def test_batch_autotune_and_constraints(oracle_mod, gpu):
    _say("case batch autotune + constraints")
    b, states = _make_batch(gpu, SHAPES, "fp64")
    rep = b.autotune(2)
    assert len(rep) >= 1
    b.step(2)
    got = b.setConstraintCalcs()
    L = oracle_mod.lib()
    for r, (shape, st) in enumerate(zip(SHAPES, states)):
        o = _oracle(oracle_mod, gpu, shape, st)
        o.run_steps(2)
        o.stepInit(); o.clearSrc(); L.or_static_add_src(o.h); L.or_lambda_add_src(o.h)
        want = o.constraints()
        for q in (0, 2, 6, 7):
            w = want["H"][q]
            assert abs(got["H"][r][q] - w) <= 1e-10 * abs(w), (r, q)
        for f in FIELDS:
            assert rel_l2(b.members[r].download(f + "_p"), o.field(f + "_p")) <= 1e-12, (r, f)


def test_batch_nan_guard_names_realisation(gpu):
    _say("case batch NaN guard")
    b, states = _make_batch(gpu, SHAPES, "fp64")
    ph = states[2]["DIFFphi"].copy()
    ph[3, 4, 5] = np.nan
    b.members[2].upload("DIFFphi_p", ph)
    with pytest.raises(gpu.CosmoGpuError) as ei:
        b.step(1)
    assert ei.value.code == gpu.ERR_NAN
    assert "2" in str(ei.value)
    # the other realisations evolved normally
    assert b.members[0].numNaNs("DIFFphi_p") == 0


def test_batch_ensemble_sizes_vs_oracle(oracle_mod, gpu):
    """One realisation of each grid size of the ensemble configuration (48^3,
    96^3, 128^3; order 4, P(k) ~ k^-2 constraint-solved ICs, rms 1e-2) in one
    batch, 2 steps, against the oracle: FP64 to 1e-12, FP32 split to 1e-5."""
    from cosmograph_amd.ics import static_dust_ics
    sizes = (48, 96, 128)
    ics = [static_dust_ics(n, -2.0, 1e-2, seed=n) for n in sizes]
    want = []
    for n, (fields, frw) in zip(sizes, ics):
        _say(f"case ensemble sizes: oracle {n}^3")
        o = oracle_mod.Sim(oracle_mod.make_params((n, n, n), order=4))
        for name, arr in fields.items():
            o.field(name)[:] = arr
        o.frw_set(frw["phi"], frw["K"])
        for rho, w in frw["fluids"]:
            o.frw_add_fluid(rho, w)
        o.run_steps(2)
        want.append({f: o.field(f + "_p").copy() for f in FIELDS})
        del o
    for prec, tol in (("fp64", 1e-12), ("fp32_split", 1e-5)):
        _say(f"case ensemble sizes: batch {prec}")
        b = gpu.Batch([gpu.GpuParams(n=n, order=4, precision=prec, matter="static") for n in sizes])
        for m, (fields, frw) in zip(b.members, ics):
            for name, arr in fields.items():
                m.upload(name, arr)
            m.frw_set(frw["phi"], frw["K"])
            for rho, w in frw["fluids"]:
                m.frw_add_fluid(rho, w)
        b.step(2)
        for n, m, w in zip(sizes, b.members, want):
            for f in FIELDS:
                err = rel_l2(m.download(f + "_p"), w[f])
                assert err <= (tol if np.any(w[f]) else 1e-30), (prec, n, f, err)
        del b


def test_batch_rejects_mixed_precision(gpu):
    with pytest.raises(gpu.CosmoGpuError):
        gpu.Batch([gpu.GpuParams(n=16), gpu.GpuParams(n=16, precision="fp32_split")])
