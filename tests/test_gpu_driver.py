"""The `cosmo_gpu` driver end to end on the GPU: a config file and initial
data in, the StaticSim / DustSim step sequence of the reference
(`steps = N` runs N+1 RK4 steps, output before each step, sims/sim.cc:101-131),
the constraint and FRW series in the reference's gzip text format, and the
final state, which must match the oracle run for the same number of steps."""
import gzip
import os
import subprocess
import sys

import numpy as np
import pytest

from common import FIELDS, rel_l2, smooth_random_state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cosmograph_amd", "cosmo_gpu")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_driver(cfg_path):
    r = subprocess.run([DRIVER, str(cfg_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def rows(path):
    with gzip.open(path, "rt") as fh:
        return [[float(x) for x in line.split("\t") if x] for line in fh.read().splitlines()]


def oracle_output_constraints(oracle_mod, o, dust):
    L = oracle_mod.lib()
    L.or_bssn_step_init(o.h)
    if dust:
        L.or_dust_step_init(o.h)
    L.or_bssn_clear_src(o.h)
    (L.or_dust_add_src if dust else L.or_static_add_src)(o.h)
    L.or_lambda_add_src(o.h)
    return o.constraints()


@pytest.mark.parametrize("kind", ["static", "dust"])
def test_driver_matches_oracle(oracle_mod, tmp_path, kind):
    from cosmograph_amd.ics import save_ics, static_dust_ics
    n, steps, lam = 20, 2, 0.01
    _say(f"case driver {kind} n={n} steps={steps}")
    if kind == "static":
        fields, frw = static_dust_ics(n, -2.0, 1e-3, 5)
    else:
        st = smooth_random_state((n, n, n), seed=71, with_alpha=False)
        fields = {f + "_p": st[f] for f in FIELDS}
        fields["dust_D_p"] = 0.05 + 0.1 * st["DIFFD"]
        fields["dust_S1_p"] = 0.5 * st["DIFFK"]
        frw = {"phi": 0.0, "K": 0.0, "fluids": []}
    save_ics(str(tmp_path / "ic"), fields, frw, lam)
    out = tmp_path / "out"
    cfg = tmp_path / "config.txt"
    cfg.write_text(f"""N = {n}
stencil_order = 4
H_LEN_FRAC = 1
dt_frac = 0.1
simulation_type = {kind}
lapse = Static
steps = {steps}
ic_dir = {tmp_path / 'ic'}
output_dir = {out}
IO_constraint_interval = 1
IO_bssnstats_interval = 1
dump_file = calculated
IO_final_state = 1
IO_1D_grid_interval = 1
IO_1D_DIFFphi_a = 1
IO_1D_DIFFK_a = 1
IO_1D_DIFFK_a_axis = 3
IO_1D_DIFFK_a_xoffset = 4
IO_1D_DIFFK_a_yoffset = 7
IO_1D_DIFFr_a = 1
IO_1D_A11_a = 0
IO_3D_grid_interval = 2
IO_3D_DIFFK_a = 1
IO_powerspec_interval = 2
""")
    r = run_driver(cfg)
    assert "RK steps: 3" in r.stdout                        # steps = 2 runs 3 RK4 steps

    # the oracle, same data and sequence
    p = oracle_mod.make_params(n, order=4, matter=kind, lam=lam)
    o = oracle_mod.Sim(p)
    for name, arr in fields.items():
        o.field(name)[:] = arr
    o.frw_set(frw["phi"], frw["K"])
    for rho, w in frw["fluids"]:
        o.frw_add_fluid(rho, w)
    want_H, want_strips, want_stats, want_3d, want_spec = [], [], [], {}, []
    for s in range(steps + 1):
        want_H.append(oracle_output_constraints(oracle_mod, o, kind == "dust")["H"])
        want_strips.append({"DIFFphi_a": o.field("DIFFphi_a")[:, 0, 0].copy(),
                            "DIFFK_a": o.field("DIFFK_a")[4, 7, :].copy(),
                            "DIFFr_a": o.field("DIFFr_a")[:, 0, 0].copy()})
        # (value, scale of the field it summarises)
        want_stats.append([(v, np.abs(o.field(name)).max()) for name in ("DIFFphi_a", "DIFFK_a", "DIFFr_a", "ricci_a")
                           for v in o.conformal_stats(name)[:2]] + [(o.conformal_stats("DIFFphi_a")[2], 1.0)])
        if s == 0:
            want_1d_H, want_1d_M = o.constraint_strip(1, 0, 0)
        if s % 2 == 0:
            want_3d[s] = o.field("DIFFK_a").copy()
            want_spec.append(oracle_mod.power_dump(o.field("DIFFphi_a")))
            want_spec.append(oracle_mod.power_dump(o.field("DIFFr_a")))
        o.run_steps(1)
    names = (out / "final_fields.txt").read_text().split()
    assert len(names) == (22 if kind == "dust" else 18)
    for name in names:
        got = np.fromfile(out / (name + ".f64"), dtype="<f8").reshape(n, n, n)
        ref = o.field(name)
        assert rel_l2(got, ref) <= (1e-12 if np.any(ref != 0) else 1e-30), name

    H = rows(out / "H_violations.dat.gz")
    assert len(H) == steps + 1 and all(len(r_) == 5 for r_ in H)
    for s in range(steps + 1):
        # columns: mean, stdev, max of H/[H], max H, L2 (IO/io.cc:249-253)
        for col, q in zip(range(5), (4, 5, 6, 2, 7)):
            w = want_H[s][q]
            assert abs(H[s][col] - w) <= 1e-9 * max(abs(w), abs(want_H[s][2]) * 1e-3), (s, col, H[s][col], w)
    for f in ("M_violations", "G_violations", "A_violations", "S_violations"):
        assert len(rows(out / (f + ".dat.gz"))) == steps + 1
    frw_rows = rows(out / "calculated.frwdat.gz")
    assert len(frw_rows) == steps + 1 and all(len(r_) == 4 for r_ in frw_rows)

    # io_bssn_fields_snapshot: one strip per step and field (IO/io.cc:140-158)
    assert not (out / "1D_A11_a.strip.dat.gz").exists()
    for name in ("DIFFphi_a", "DIFFK_a", "DIFFr_a"):
        got = rows(out / f"1D_{name}.strip.dat.gz")
        assert len(got) == steps + 1, name
        for s in range(steps + 1):
            w = want_strips[s][name]
            assert len(got[s]) == w.size, name
            np.testing.assert_allclose(got[s], w, rtol=0, atol=1e-12 * max(np.abs(w).max(), 1e-300), err_msg=name)
    for s, w in want_3d.items():
        got = np.fromfile(out / f"3D_DIFFK_a{s}.3d_grid.f64", dtype="<f8").reshape(n, n, n)
        assert rel_l2(got, w) <= 1e-12
    assert not (out / "3D_DIFFK_a1.3d_grid.f64").exists()
    # io_bssn_dump_statistics: conformal average / stdev of DIFFphi, DIFFK,
    # DIFFr, ricci and the volume average (IO/io.cc:386-403)
    st = rows(out / "calculated.dat.gz")
    assert len(st) == steps + 1 and all(len(r_) == 9 for r_ in st)
    for s in range(steps + 1):
        for col in range(9):
            w, scale = want_stats[s][col]
            assert abs(st[s][col] - w) <= 1e-10 * abs(w) + 1e-13 * scale, (s, col, st[s][col], w)
    # io_bssn_constraint_violation's 1D strips, at step 0 only (the default
    # IO_constraint_snapshot_interval, IO/io.cc:186-235)
    h1 = rows(out / "1D_hamiltonian_constraints.strip.dat.gz")
    m1 = rows(out / "1D_momentum_constraints.strip.dat.gz")
    assert len(h1) == 1 and len(m1) == 1 and len(h1[0]) == n
    np.testing.assert_allclose(h1[0], want_1d_H, rtol=0, atol=1e-10 * np.abs(want_1d_H).max())
    np.testing.assert_allclose(m1[0], want_1d_M, rtol=0, atol=1e-10 * np.abs(want_1d_M).max())
    # io_bssn_fields_powerdump: DIFFphi_a then DIFFr_a at every second step, "%g" text
    spec = rows(out / "spec.dat.gz")
    assert len(spec) == len(want_spec)
    for got, w in zip(spec, want_spec):
        np.testing.assert_allclose(got, w, rtol=2e-5, atol=1e-12 * max(w.max(), 1e-300))


def test_driver_bardeen_svt_output(oracle_mod, tmp_path):
    """use_bardeen in a dust run (no FRW reference, as the reference requires):
    Bardeen::setPotentials at every output step, SVT_violations.dat.gz every
    SVT_constraint_interval steps (17 tab-separated values a row,
    IO/io.cc:1116-1131), and the Bardeen_Phi field in the 3D dumps."""
    import bardeen as bd
    from cosmograph_amd.ics import save_ics
    n, steps = 20, 1
    _say(f"case driver bardeen n={n} steps={steps}")
    st = smooth_random_state((n, n, n), seed=73)
    st["DIFFK"] += -0.05
    fields = {f + "_p": st[f] for f in FIELDS}
    fields["dust_D_p"] = 0.05 + 0.1 * st["DIFFD"]
    fields["dust_S2_p"] = 0.3 * st["DIFFK"]
    frw = {"phi": 0.0, "K": 0.0, "fluids": []}
    save_ics(str(tmp_path / "ic"), fields, frw, 0.0)
    out = tmp_path / "out"
    cfg = tmp_path / "config.txt"
    cfg.write_text(f"""N = {n}
stencil_order = 4
H_LEN_FRAC = 1
simulation_type = dust
lapse = Harmonic
steps = {steps}
ic_dir = {tmp_path / 'ic'}
output_dir = {out}
use_bardeen = 1
Omega_L = 0.2
SVT_constraint_interval = 1
IO_3D_grid_interval = 1
IO_3D_Bardeen_Phi = 1
""")
    run_driver(cfg)
    got = rows(out / "SVT_violations.dat.gz")
    assert len(got) == steps + 1 and all(len(r_) == 17 for r_ in got)
    o = oracle_mod.Sim(oracle_mod.make_params(n, order=4, matter="dust", lapse="Harmonic"))
    for name, arr in fields.items():
        o.field(name)[:] = arr
    L = oracle_mod.lib()
    dt = o.p.dt
    for s in range(steps + 1):
        L.or_bssn_step_init(o.h)
        L.or_dust_step_init(o.h)
        L.or_bssn_clear_src(o.h)
        L.or_dust_add_src(o.h)
        L.or_lambda_add_src(o.h)
        o.RKEvolve()
        a_reg = {f: o.field(f + "_a").copy() for f in FIELDS}
        c_reg = {f: o.field(f + "_c").copy() for f in FIELDS}
        wf, wv, _ = bd.set_potentials(a_reg, c_reg, s * dt, 4, 1.0 / n, 1.0, use_ml=True, omega_L=0.2)
        for q in range(17):
            # %.15g text: 15 significant digits
            assert abs(got[s][q] - wv[q]) <= 1e-9 * abs(wv[q]) + 1e-300, (s, q, got[s][q], wv[q])
        phi3d = np.fromfile(out / f"3D_Bardeen_Phi{s}.3d_grid.f64", dtype="<f8").reshape(n, n, n)
        assert rel_l2(phi3d, wf["Phi"]) <= 1e-10
        o.run_steps(1)
