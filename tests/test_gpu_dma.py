"""LDS-DMA halo staging against register staging.

Both staging paths put the same values into LDS and differentiate them with
the same source code; the two kernels are compiled separately, so the
compiler may contract a few multiply-adds differently.  After whole RK4 steps
every register must agree to rounding: relative L2 <= 1e-15 in FP64 and
<= 1e-7 in FP32 split (a wrong halo element gives errors many orders of
magnitude larger).  The DMA cases are those the library uses by default
(nz % 4 == 0, stages without metric normalisation); kernel_info proves the
DMA kernel is the one selected (its LDS layout differs).  Parity of the
default path with the oracle is covered by the other GPU tests, which run
with DMA on."""
import sys

import numpy as np
import pytest

from common import FIELDS, smooth_random_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import cosmograph_amd as cg
    cg.load_library()
    return cg


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _load(g, shape, seed, lapse):
    st = smooth_random_state(shape, seed=seed, with_alpha=lapse != "Static")
    for f in FIELDS + ["DIFFD"]:
        g.upload(f + ("_a" if f == "DIFFD" else "_p"), st[f])
    g.frw_set(0.0, -3.0)
    g.frw_add_fluid(3.0 / (8.0 * np.pi), 0.0)


def _compare(got, ref, prec):
    worst = 0.0
    for f in FIELDS:
        d = np.abs(got[f] - ref[f])
        n = np.sqrt(np.sum(ref[f] ** 2))
        rel = np.sqrt(np.sum(d ** 2)) / n if n > 0 else np.sqrt(np.sum(d ** 2))
        if d.max() > 0:
            _say(f"  {f}: rel_l2 {rel:.3e} max_abs {d.max():.3e} points {int(np.count_nonzero(d))}/{d.size}")
        worst = max(worst, rel)
    _say(f"  worst rel_l2 {worst:.3e}")
    assert worst <= (1e-15 if prec == "fp64" else 1e-7)


CASES = [
    # precision, order, shape, lapse, normalize_metric
    ("fp32_split", 6, (32, 32, 32), "Static", 1),
    ("fp32_split", 4, (20, 24, 28), "Harmonic", 1),
    ("fp32_split", 6, (24, 20, 36), "Static", 0),
    ("fp64", 6, (24, 20, 36), "Static", 0),
    ("fp64", 4, (16, 16, 20), "Harmonic", 1),
    ("fp64", 6, (18, 22, 40), "Harmonic", 1),
]


@pytest.mark.parametrize("prec,order,shape,lapse,norm", CASES)
def test_dma_matches_register_staging(gpu, prec, order, shape, lapse, norm):
    _say(f"case dma {prec} order={order} shape={shape} lapse={lapse} norm={norm}")
    out, info = {}, {}
    for dma in (False, True):
        g = gpu.BSSN(gpu.GpuParams(n=shape, order=order, precision=prec, lapse=lapse, normalize_metric=norm))
        g.set_dma(dma)
        _load(g, shape, seed=order + len(lapse), lapse=lapse)
        g.step(2)
        out[dma] = {f: g.download(f + "_p") for f in FIELDS}
        info[dma] = g.kernel_info()
        del g
    assert info[True]["lds_bytes"] != info[False]["lds_bytes"], info  # the DMA twin ran
    _compare(out[True], out[False], prec)


def test_register_path_when_nz_not_multiple_of_4(gpu):
    shape = (16, 16, 18)
    _say(f"case dma fallback shape={shape}")
    g = gpu.BSSN(gpu.GpuParams(n=shape, order=4, precision="fp32_split"))
    h = gpu.BSSN(gpu.GpuParams(n=shape, order=4, precision="fp32_split"))
    h.set_dma(False)
    assert g.kernel_info()["lds_bytes"] == h.kernel_info()["lds_bytes"]


def test_batch_dma_matches_register_staging(gpu):
    shapes = [(16, 16, 16), (24, 24, 24), (20, 16, 32)]
    _say(f"case batch dma shapes={shapes}")
    out = {}
    for dma in (False, True):
        b = gpu.Batch([gpu.GpuParams(n=s, order=4, precision="fp32_split") for s in shapes])
        b.set_dma(dma)
        for r, (m, s) in enumerate(zip(b.members, shapes)):
            _load(m, s, seed=40 + r, lapse="Static")
        b.step(2)
        out[dma] = [{f: m.download(f + "_p") for f in FIELDS} for m in b.members]
        del b
    for r in range(len(shapes)):
        _compare(out[True][r], out[False][r], "fp32_split")


@pytest.mark.parametrize("prec,shape", [("fp32_split", (32, 32, 48)), ("fp64", (32, 32, 48)),
                                        ("fp32_split", (32, 40, 48))])
def test_tile_orders_give_identical_results(gpu, prec, shape):
    """Every tile is computed by the same code whatever the traversal order
    (0-2: one axis slowest; 3: Morton columns, here 8 x 4 tile columns; 4, 5:
    strips of 2 or 4 tile rows along j, the last one partial when the 5 tile
    rows of the second shape do not fill it)."""
    _say(f"case tile orders {prec} shape={shape}")
    out = []
    for order in range(6):
        g = gpu.BSSN(gpu.GpuParams(n=shape, order=6, precision=prec, tile=(4, 8, 8)))
        g.set_tile_order(order)
        _load(g, shape, seed=5, lapse="Static")
        g.step(1)
        out.append({f: g.download(f + "_p") for f in FIELDS})
        del g
    for order in range(1, 6):
        for f in FIELDS:
            assert np.array_equal(out[order][f], out[0][f]), (order, f)


DUST_CASES = [
    # precision, order, shape, lambda
    ("fp32_split", 6, (24, 20, 32), 0.02),
    ("fp32_split", 4, (20, 24, 16), 0.0),
    ("fp64", 6, (20, 24, 28), 0.02),
]


@pytest.mark.parametrize("prec,order,shape,lam", DUST_CASES)
def test_dust_dma_matches_register_staging(gpu, prec, order, shape, lam):
    """The dust-fluid stage kernel's DMA twin stages the 12 fluxes through
    LDS with the other FP32 (or FP64) fields."""
    _say(f"case dust dma {prec} order={order} shape={shape} lam={lam}")
    names = FIELDS + ["dust_D", "dust_S1", "dust_S2", "dust_S3"]
    out, info = {}, {}
    x = [np.arange(n) / n for n in shape]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    for dma in (False, True):
        g = gpu.BSSN(gpu.GpuParams(n=shape, order=order, precision=prec, matter="dust", lambda_=lam))
        g.set_dma(dma)
        st = smooth_random_state(shape, seed=order + 7, with_alpha=False)
        for f in FIELDS:
            g.upload(f + "_p", st[f] - (0.5 if f == "DIFFK" else 0.0))
        g.upload("dust_D_p", 0.05 * (1.0 + 0.02 * np.cos(2 * np.pi * (X + Y))))
        g.upload("dust_S1_p", 1e-3 * np.sin(2 * np.pi * (Z - 2 * X)))
        g.upload("dust_S3_p", 1e-3 * np.cos(2 * np.pi * (Y + Z)))
        g.step(2)
        out[dma] = {f: g.download(f + "_p") for f in names}
        info[dma] = g.kernel_info()
        del g
    assert info[True]["lds_bytes"] != info[False]["lds_bytes"], info  # the DMA twin ran
    worst = 0.0
    for f in names:
        ref = out[False][f]
        if not np.any(ref):
            assert not np.any(out[True][f]), f
            continue
        err = np.sqrt(np.sum((out[True][f] - ref) ** 2)) / np.sqrt(np.sum(ref ** 2))
        worst = max(worst, err)
    _say(f"  worst rel_l2 {worst:.3e}")
    assert worst <= (1e-15 if prec == "fp64" else 1e-7)
