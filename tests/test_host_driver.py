"""The C++ host classes and the `cosmo_gpu` driver (cosmograph_amd/host):
config keys resolve as the reference's do, errors exit with the reference's
throw codes.  Nothing here touches a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cosmograph_amd", "cosmo_gpu")


def run(args, **kw):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} not built (run __graft_entry__.build())")
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=60, **kw)


def params(tmp_path, text):
    cfg = tmp_path / "config.txt"
    cfg.write_text(text)
    r = run(["--print-params", str(cfg)])
    assert r.returncode == 0, r.stderr
    return dict(kv.split("=") for kv in r.stdout.split())


def test_defaults_follow_the_reference(tmp_path):
    p = params(tmp_path, "N = 32\nsteps = 10\nsimulation_type = static\n")
    assert (int(p["nx"]), int(p["ny"]), int(p["nz"])) == (32, 32, 32)
    assert int(p["stencil_order"]) == 8            # STENCIL_ORDER default (cosmo_macros.h)
    assert float(p["dx"]) == 0.5 / 32              # H_LEN_FRAC = 0.5
    assert float(p["dt"]) == 0.1 * (0.5 / 32)      # dt_frac = 0.1 (cosmo.cc:42)
    assert int(p["normalize_metric"]) == 1 and float(p["ko_coeff"]) == 0.0
    assert int(p["lapse"]) == 0 and int(p["matter"]) == 1 and int(p["precision"]) == 0


def test_keys_map_onto_parameters(tmp_path):
    p = params(tmp_path, """NX = 16
NY = 24
NZ = 32
stencil_order = 4
H_LEN_FRAC = 1
dt_frac = 0.2
KO_damping_coefficient = 0.05
a_adj_amp = 0.1
k_damping_amp = 0.3
normalize_metric = 0
lapse = ConformalFLRW
simulation_type = dust
gpu_precision = fp64
gpu_device = 0
steps = 3
""")
    assert (int(p["nx"]), int(p["ny"]), int(p["nz"])) == (16, 24, 32)
    assert int(p["stencil_order"]) == 4 and float(p["dx"]) == 1.0 / 16
    assert float(p["dt"]) == 0.2 * (1.0 / 16)
    assert float(p["ko_coeff"]) == 0.05 and float(p["a_adj_amp"]) == 0.1 and float(p["k_damping_amp"]) == 0.3
    assert int(p["normalize_metric"]) == 0 and int(p["lapse"]) == 2 and int(p["matter"]) == 2


@pytest.mark.parametrize("text,code", [
    ("N = 16\nlapse = GeneralizedNewton\nsteps = 1\n", 1),  # not on this path: throw -1
    ("N = 16\nshift = GammaDriver\nsteps = 1\n", 1),
    ("steps = 1\n", 1),                                  # N is required: throw -1
])
def test_rejected_configurations(tmp_path, text, code):
    cfg = tmp_path / "config.txt"
    cfg.write_text(text)
    r = run(["--print-params", str(cfg)])
    assert r.returncode == code


def test_unknown_simulation_type_throws_2(tmp_path):
    cfg = tmp_path / "config.txt"
    cfg.write_text(f"N = 16\nsteps = 1\nsimulation_type = particles\noutput_dir = {tmp_path / 'out'}\n")
    assert run([str(cfg)]).returncode == 2     # cosmo.cc:88-92


def test_missing_config_file():
    assert run(["/nonexistent/config.txt"]).returncode == 1


def test_save_ics_layout(tmp_path):
    import numpy as np
    from cosmograph_amd.ics import save_ics
    a = np.arange(8.0).reshape(2, 2, 2)
    # numpy scalars (as the IC generator returns them) must be written as plain numbers
    save_ics(str(tmp_path / "ic"), {"DIFFphi_p": a},
             {"phi": 0, "K": np.float64(-3.0), "fluids": [(np.float64(0.1), 0.0)]}, 0.5)
    assert (tmp_path / "ic" / "fields.txt").read_text().split() == ["DIFFphi_p"]
    assert np.array_equal(np.fromfile(tmp_path / "ic" / "DIFFphi_p.f64", dtype="<f8"), a.ravel())
    lines = (tmp_path / "ic" / "frw.txt").read_text().splitlines()
    assert lines[0].split() == ["0.0", "-3.0"] and lines[1].split()[0] == "fluid" and lines[2] == "lambda 0.5"
