"""ctypes binding for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product package
(cosmograph_amd) never imports this module.

The oracle restates CosmoGRaPH's BSSN hot path in C (oracle/oracle.c); the
Sim class below mirrors the reference's method names (stepInit, RKEvolve,
K{n}Finalize, setConstraintCalcs) so tests read like the reference's drivers.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# the same source built with the reference's optimisation flags (-O3
# -ffast-math, cmake/optimization.cmake:15-18): used only to time the CPU
# baseline, never as the checker
PERF_LIB_PATH = os.path.join(HERE, "liboracle_perf.so")
REF_DIR = os.path.join(HERE, "_ref")

NF, NSRC, NX = 18, 11, 7
FIELDS = ["DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23",
          "DIFFgamma33", "DIFFphi", "A11", "A12", "A13", "A22", "A23", "A33", "DIFFK",
          "Gamma1", "Gamma2", "Gamma3", "DIFFalpha"]
SOURCES = ["DIFFr", "DIFFS", "S1", "S2", "S3", "STF11", "STF12", "STF13", "STF22", "STF23", "STF33"]
MATTER = {"none": 0, "static": 1, "dust": 2}
LAPSE = {"Static": 0, "Harmonic": 1, "ConformalFLRW": 2, "Anharmonic": 3, "OnePlusLog": 4,
         "ConformalSync": 5, "AwAGaugeWave": 6, "AwAShiftedWave": 7, "DampedWave": 8,
         "TestAijDriverLapse": 9, "TestKDriverLapse": 10}
# shift gauges; "none" is the default build without USE_BSSN_SHIFT
SHIFT = {"none": 0, "Static": 1, "GammaDriver": 2, "DampedWave": 3, "AwAShiftedWave": 4, "AijDriverShift": 5}
# fields of the USE_BSSN_SHIFT build (auxB only with the gamma driver)
X_FIELDS = ["beta1", "beta2", "beta3", "expN", "auxB1", "auxB2", "auxB3"]


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("order", C.c_int),
                ("dx", C.c_double), ("dt", C.c_double), ("ko_coeff", C.c_double),
                ("a_adj_amp", C.c_double), ("k_damping_amp", C.c_double),
                ("normalize_metric", C.c_int), ("lapse", C.c_int), ("matter", C.c_int),
                ("lambda_", C.c_double), ("gd_c", C.c_double), ("dw_mu_l", C.c_double),
                ("dw_p", C.c_double), ("shift", C.c_int), ("gauge_wave_dir", C.c_int),
                ("gd_eta", C.c_double), ("dw_mu_s", C.c_double), ("k_driver_coeff", C.c_double)]


def _point_fields():
    d = C.c_double
    return [("v", d * NF), ("src", d * NSRC), ("dg", (d * 6) * 3), ("ddg", (d * 6) * 6),
            ("dphi", d * 3), ("ddphi", d * 6), ("dalpha", d * 3), ("ddalpha", d * 6),
            ("dK", d * 3), ("dGam", (d * 3) * 3), ("dA", (d * 6) * 3), ("ko", d * NF),
            ("phi_F", d), ("K_F", d), ("rho_F", d), ("S_F", d),
            ("K_avg", d), ("rho_avg", d), ("avg_vol", d),
            ("phi", d), ("K", d), ("r", d), ("S", d), ("alpha", d),
            ("g", d * 6), ("gi", d * 6), ("Acont", d * 6), ("AijAij", d),
            ("G", (d * 6) * 3), ("GL", (d * 6) * 3), ("Gammad", d * 3), ("DDphi", d * 6),
            ("DDaTF", d * 6), ("DDaTR", d), ("Uricci", d * 6), ("ricci_ij", d * 6),
            ("ricci", d), ("ricciTF", d * 6), ("H", d), ("rhs", d * NF),
            ("xv", d * NX), ("dbeta", (d * 3) * 3), ("ddbeta", (d * 6) * 3), ("dexpN", d * 3),
            ("adv", d * NF), ("kox", d * NX), ("rhsx", d * NX)]


class Point(C.Structure):
    _fields_ = _point_fields()


_libs = {}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib(perf: bool = False):
    """The checker library, or with perf=True its -O3 -ffast-math build
    (timing only)."""
    key = "perf" if perf else "check"
    if key not in _libs:
        path = PERF_LIB_PATH if perf else LIB_PATH
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(Params)]
        L.or_destroy.argtypes = [vp]
        L.or_field.restype = dp
        L.or_field.argtypes = [vp, C.c_char_p]
        L.or_points.restype = C.c_long
        L.or_points.argtypes = [vp]
        L.or_num_threads.restype = C.c_int
        for name in ["or_bssn_step_init", "or_bssn_rk_evolve", "or_bssn_clear_src",
                     "or_bssn_normalize", "or_bssn_extra_field_data", "or_bssn_prep_output",
                     "or_static_add_src", "or_lambda_add_src", "or_dust_step_init",
                     "or_dust_rk_evolve", "or_dust_add_src"]:
            getattr(L, name).argtypes = [vp]
        L.or_bssn_k_finalize.argtypes = [vp, C.c_int]
        L.or_dust_k_finalize.argtypes = [vp, C.c_int]
        L.or_run_steps.argtypes = [vp, C.c_int]
        L.or_run_steps.restype = C.c_int
        L.or_frw_set.argtypes = [vp, C.c_double, C.c_double]
        L.or_frw_add_fluid.argtypes = [vp, C.c_double, C.c_double]
        L.or_frw_get.argtypes = [vp, dp]
        L.or_frw_step_stage.argtypes = [vp, C.c_int, C.c_double]
        L.or_bssn_get_extra.argtypes = [vp, dp]
        L.or_bssn_constraints.argtypes = [vp, dp, dp, dp, dp, dp]
        L.or_num_nans.argtypes = [vp, C.c_char_p]
        L.or_num_nans.restype = C.c_long
        L.or_point_compute.argtypes = [C.POINTER(Params), C.POINTER(Point)]
        L.or_point_momentum.argtypes = [C.POINTER(Point), dp]
        L.or_point_constraints.argtypes = [C.POINTER(Point), dp]
        L.or_constraint_points.argtypes = [vp, dp]
        L.or_stencil.argtypes = [vp, dp, C.c_int, C.c_int, C.c_int, C.c_long, C.c_long, C.c_long]
        L.or_stencil.restype = C.c_double
        L.or_reduction.argtypes = [vp, C.c_int, dp, dp, C.c_double]
        L.or_reduction.restype = C.c_double
        L.or_constraint_strip.argtypes = [vp, C.c_int, C.c_long, C.c_long, dp, dp]
        L.or_rk4_demo.argtypes = [dp, C.c_long, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]
        L.or_upwind.argtypes = [vp, dp, dp, dp]
        _libs[key] = L
    return _libs[key]


def make_params(n, order=4, dx=None, dt_frac=0.1, ko=0.0, a_adj=0.0, k_damp=0.0,
                normalize=1, lapse="Static", matter="static", lam=0.0, box=1.0, gd_c=1.0,
                dw_mu_l=0.0, dw_p=0.0, shift="none", gauge_wave_dir=1, gd_eta=0.0, dw_mu_s=0.0,
                k_driver_coeff=0.04):
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    dx = box / nx if dx is None else dx
    return Params(nx, ny, nz, order, dx, dt_frac * dx, ko, a_adj, k_damp, normalize,
                  LAPSE[lapse], MATTER[matter], lam, gd_c, dw_mu_l, dw_p, SHIFT[shift], gauge_wave_dir,
                  gd_eta, dw_mu_s, k_driver_coeff)


class Sim:
    """Oracle simulation state; method names follow components/bssn/bssn.h."""

    def __init__(self, params: Params, perf: bool = False):
        self.p = params
        self.L = lib(perf)
        self.h = self.L.or_create(C.byref(params))
        self.shape = (params.nx, params.ny, params.nz)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_destroy(self.h)
            self.h = None

    def field(self, name) -> np.ndarray:
        ptr = self.L.or_field(self.h, name.encode())
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=self.shape)

    # FRW
    def frw_set(self, phi, K):
        self.L.or_frw_set(self.h, phi, K)

    def frw_add_fluid(self, rho, w):
        self.L.or_frw_add_fluid(self.h, rho, w)

    def frw_get(self):
        out = (C.c_double * 4)()
        self.L.or_frw_get(self.h, out)
        return tuple(out)

    # BSSN methods
    def stepInit(self):
        self.L.or_bssn_step_init(self.h)

    def RKEvolve(self):
        self.L.or_bssn_rk_evolve(self.h)

    def KFinalize(self, n):
        self.L.or_bssn_k_finalize(self.h, n)

    def clearSrc(self):
        self.L.or_bssn_clear_src(self.h)

    def prepOutput(self):
        self.L.or_bssn_prep_output(self.h)

    def extra(self):
        out = (C.c_double * 4)()
        self.L.or_bssn_get_extra(self.h, out)
        return dict(zip(["K_min", "K_avg", "avg_vol", "rho_avg"], out))

    def run_steps(self, n):
        rc = self.L.or_run_steps(self.h, n)
        if rc:
            raise RuntimeError(f"oracle run failed with code {rc} (10 = NaN)")

    def constraints(self):
        H, M = (C.c_double * 8)(), (C.c_double * 8)()
        G, A, S = (C.c_double * 7)(), (C.c_double * 7)(), (C.c_double * 7)()
        self.L.or_bssn_constraints(self.h, H, M, G, A, S)
        return {"H": list(H), "M": list(M), "G": list(G), "A": list(A), "S": list(S)}

    def constraint_points(self) -> np.ndarray:
        """[nx, ny, nz, 5, 3]: {val, scale, scaled} of H, M, G, A, S at every
        point of _a (what setConstraintCalcs reduces)."""
        out = np.zeros(self.shape + (5, 3))
        self.L.or_constraint_points(self.h, dptr(out))
        return out

    def num_nans(self, name="DIFFphi_a"):
        return self.L.or_num_nans(self.h, name.encode())

    def constraint_strip(self, axis=1, n1=0, n2=0):
        n = (self.p.nx, self.p.ny, self.p.nz)[axis - 1]
        H, M = np.zeros(n), np.zeros(n)
        self.L.or_constraint_strip(self.h, axis, n1, n2, dptr(H), dptr(M))
        return H, M

    def conformal_stats(self, name):
        """conformal_average, conformal_standard_deviation of a field and
        volume_average of DIFFphi_a, with the current FRW phi."""
        phiF = self.frw_get()[0]
        f, ph = dptr(self.field(name)), dptr(self.field("DIFFphi_a"))
        return tuple(self.L.or_reduction(self.h, k, f, ph, phiF) for k in (2, 6, 1))


def power_dump(field: np.ndarray) -> np.ndarray:
    """Fourier::powerDump (utils/Fourier.h:147-220): the r2c transform of the
    field, |F|^2 binned by floor(|p|) with p in lattice units, each mode with
    0 < pz < NZ/2 counted twice (it stands for its conjugate too) and pz = 0,
    NZ/2 once; the bin averages.  numpy's rfftn stands in for FFTW's r2c."""
    nx, ny, nz = field.shape
    nbins = int(np.sqrt(float(nx * nx + ny * ny + nz * nz)) / 2.0) + 1
    F = np.fft.rfftn(np.asarray(field, dtype=np.float64))
    px = np.where(np.arange(nx) <= nx // 2, np.arange(nx), np.arange(nx) - nx)
    py = np.where(np.arange(ny) <= ny // 2, np.arange(ny), np.arange(ny) - ny)
    pz = np.arange(nz // 2 + 1)
    PX, PY, PZ = np.meshgrid(px, py, pz, indexing="ij")
    bins = np.sqrt((PX * PX + PY * PY + PZ * PZ).astype(np.float64)).astype(np.int64)
    w = np.where((PZ >= 1) & (PZ < nz // 2), 2, 1)
    f2 = np.bincount(bins.ravel(), weights=(w * np.abs(F) ** 2).ravel(), minlength=nbins)
    cnt = np.bincount(bins.ravel(), weights=w.ravel(), minlength=nbins)
    return np.where(cnt > 0, f2 / np.maximum(cnt, 1), 0.0)


def upwind(sim: Sim, f: np.ndarray, c: np.ndarray) -> np.ndarray:
    """upwind_derivative (utils/math.h:1247-1252) of f along each axis with
    the per-point speed c: [3, nx, ny, nz]."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    out = np.zeros((3,) + f.shape)
    sim.L.or_upwind(sim.h, dptr(f), dptr(c), dptr(out))
    return out


def point_compute(params: Params, pt: Point) -> Point:
    lib().or_point_compute(C.byref(params), C.byref(pt))
    return pt


def point_momentum(pt: Point):
    out = (C.c_double * 3)()
    lib().or_point_momentum(C.byref(pt), out)
    return np.array(out[:])


def point_constraints(pt: Point) -> np.ndarray:
    """[5, 3]: {val, scale, scaled} of H, M, G, A, S at a point filled by
    point_compute."""
    out = (C.c_double * 15)()
    lib().or_point_constraints(C.byref(pt), out)
    return np.array(out[:]).reshape(5, 3)


# --------------------------------------------------------------- reference
def ref_available(order: int) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"librefdrv_o{order}.so"))


def ref_lib(order: int):
    L = C.CDLL(os.path.join(REF_DIR, f"librefdrv_o{order}.so"))
    dp = C.POINTER(C.c_double)
    L.ref_info.argtypes = [C.POINTER(C.c_int)]
    L.ref_set_steps.argtypes = [C.c_double, C.c_double]
    L.ref_derivatives.argtypes = [dp, dp]
    L.ref_double_derivatives.argtypes = [dp, dp]
    L.ref_ko.argtypes = [dp, C.c_double, dp]
    L.ref_upwind.argtypes = [dp, dp, dp]
    L.ref_reductions.argtypes = [dp, dp, C.c_double, dp]
    L.ref_static_src.argtypes = [dp, dp, C.c_double, C.c_double, dp]
    L.ref_rk4.argtypes = [dp, C.c_long, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]
    L.ref_frw.argtypes = [C.c_double, C.c_double, C.c_int, dp, dp, C.c_double, C.c_int, dp]
    return L


def ref_dims(L):
    out = (C.c_int * 4)()
    L.ref_info(out)
    return tuple(out)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))
