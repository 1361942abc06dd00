"""C ABI checks that need no GPU: the built library exports every entry point
include/*.h declares, the ctypes front end binds each of them, and argument
validation rejects bad parameters before any device call."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cg_\w+)\s*\(", re.M)


def declared():
    names = []
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names += DECL.findall(txt)
    return names


@pytest.fixture(scope="module")
def lib():
    import cosmograph_amd as cg
    path = cg.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    return C.CDLL(path)


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 20
    assert len(set(names)) == len(names)
    for must in ("cg_create", "cg_rk4_steps", "cg_constraints", "cg_extra_field_data", "cg_num_nans"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_front_end_binds_every_symbol():
    import cosmograph_amd.gpu as gpu
    src = open(gpu.__file__).read()
    missing = [n for n in declared() if f'"{n}"' not in src]
    assert not missing, missing


@pytest.mark.parametrize("bad", [
    {"n": (16, 16, 0), "dx": 1.0 / 16}, {"n": (2048, 1024, 1024)},
    # 32-bit byte offsets of the FP64 registers: 1024*1024*520 points > 2^29 (split precision too)
    {"n": (1024, 1024, 520), "precision": "fp32_split"}, {"n": (1024, 1024, 513)}, {"order": 3}, {"precision": "bf16"}, {"lapse": "GeneralizedNewton"}, {"dt_frac": float("nan")},
    # the shift build carries BSSN + static dust only
    {"shift": "Static", "matter": "dust"}, {"shift": "Shiftless"},
    # all-FP32 storage: BSSN + static dust on single grids, the stage kernel's own gauges
    {"precision": "fp32_all", "matter": "dust"}, {"precision": "fp32_all", "shift": "Static"},
    {"precision": "fp32_all", "lapse": "TestKDriverLapse"},
    {"precision": "fp32_fix", "shift": "GammaDriver"}, {"precision": "fp32_fix", "lapse": "TestKDriverLapse"},
])
def test_create_rejects_bad_parameters(bad):
    import cosmograph_amd as cg
    kw = {"n": 16, "order": 4}
    kw.update(bad)
    try:
        p = cg.GpuParams(**kw)
        p.to_c()
    except KeyError:
        return  # rejected by the front end's name tables
    with pytest.raises(cg.CosmoGpuError) as ei:
        cg.BSSN(p)
    assert ei.value.code == -1


def test_null_context_is_an_argument_error(lib):
    lib.cg_sync.argtypes = [C.c_void_p]
    lib.cg_sync.restype = C.c_int
    assert lib.cg_sync(None) == -1
    lib.cg_last_error.argtypes = [C.c_void_p]
    lib.cg_last_error.restype = C.c_char_p
    assert lib.cg_last_error(None) == b"null context"
