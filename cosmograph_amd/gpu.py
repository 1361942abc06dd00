"""ctypes front end of libcosmo_gpu.so (C ABI: include/cosmo_gpu.h).

``BSSN`` mirrors the public methods of the reference's ``cosmo::BSSN``
(components/bssn/bssn.h:53-170) that its sims call: ``stepInit``,
``RKEvolve``, ``K1Finalize``..``K4Finalize``, ``step`` and the ``fields``
map, with the FRW reference integrator (``frw``) and Lambda attached as in
the reference.  Errors raise ``CosmoGpuError`` carrying the library's
message, in place of the reference's integer ``throw``s.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

FIELDS = ["DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23",
          "DIFFgamma33", "DIFFphi", "A11", "A12", "A13", "A22", "A23", "A33", "DIFFK",
          "Gamma1", "Gamma2", "Gamma3", "DIFFalpha"]
PRECISION = {"fp64": 0, "fp32_split": 1, "fp32_all": 2, "fp32_fix": 3}
MATTER = {"none": 0, "vacuum": 0, "static": 1, "dust": 2}
LAPSE = {"": 0, "Static": 0, "Harmonic": 1, "ConformalFLRW": 2, "Anharmonic": 3, "OnePlusLog": 4,
         "ConformalSync": 5, "AwAGaugeWave": 6, "AwAShiftedWave": 7, "DampedWave": 8, "TestAijDriverLapse": 9,
         "TestKDriverLapse": 10}
# shift gauges (BSSNGaugeHandler.h:113-135); "none" is the reference's default
# build without USE_BSSN_SHIFT, any other name the shift build with that gauge
SHIFT = {"none": 0, "Static": 1, "GammaDriver": 2, "DampedWave": 3, "AwAShiftedWave": 4, "AijDriverShift": 5}
X_FIELDS = ["beta1", "beta2", "beta3", "expN", "auxB1", "auxB2", "auxB3"]
ERR_NAN = -10


class CosmoGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[cosmo_gpu {code}] {msg}")
        self.code = code


class _Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("stencil_order", C.c_int),
                ("precision", C.c_int), ("dx", C.c_double), ("dt", C.c_double),
                ("ko_coeff", C.c_double), ("a_adj_amp", C.c_double),
                ("k_damping_amp", C.c_double), ("normalize_metric", C.c_int),
                ("lapse", C.c_int), ("matter", C.c_int), ("lambda_", C.c_double),
                ("device", C.c_int), ("tile_i", C.c_int), ("tile_j", C.c_int),
                ("tile_k", C.c_int), ("shift", C.c_int)]


@dataclass
class GpuParams:
    """Grid and physics parameters (reference macros and config keys)."""
    n: int | tuple = 32
    order: int = 4
    precision: str = "fp64"
    box: float = 1.0            # H_LEN_FRAC
    dx: float | None = None
    dt_frac: float = 0.1
    ko_coeff: float = 0.0
    a_adj_amp: float = 0.0
    k_damping_amp: float = 0.0
    normalize_metric: int = 1
    lapse: str = "Static"
    matter: str = "static"
    lambda_: float = 0.0
    device: int = 0
    tile: tuple | None = None
    gd_c: float = 1.0           # gauge parameters (BSSNGaugeHandler.h:139-149)
    dw_mu_l: float = 0.0
    dw_p: float = 0.0
    shift: str = "none"         # shift gauge of the shift build ("none": default build)
    gauge_wave_dir: int = 1     # BSSNGaugeHandler.h:139-149, bssn.cc:18
    gd_eta: float = 0.0
    dw_mu_s: float = 0.0
    k_driver_coeff: float = 0.04

    def shape(self):
        return (self.n, self.n, self.n) if np.isscalar(self.n) else tuple(self.n)

    def to_c(self) -> _Params:
        nx, ny, nz = self.shape()
        dx = self.box / nx if self.dx is None else self.dx
        ti, tj, tk = self.tile if self.tile else (0, 0, 0)
        return _Params(nx, ny, nz, self.order, PRECISION[self.precision], dx, self.dt_frac * dx,
                       self.ko_coeff, self.a_adj_amp, self.k_damping_amp, self.normalize_metric,
                       LAPSE[self.lapse], MATTER[self.matter], self.lambda_, self.device,
                       ti, tj, tk, SHIFT[self.shift])


_lib = None


def library_path() -> str:
    return os.path.join(HERE, "libcosmo_gpu.so")


def load_library():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise CosmoGpuError(-1, f"{path} is missing: build it with `python -m cosmograph_amd.build`")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "cg_create": ([C.POINTER(_Params), C.POINTER(vp)], C.c_int),
        "cg_destroy": ([vp], C.c_int),
        "cg_last_error": ([vp], C.c_char_p),
        "cg_field_upload": ([vp, C.c_char_p, dp, C.c_size_t], C.c_int),
        "cg_field_download": ([vp, C.c_char_p, dp, C.c_size_t], C.c_int),
        "cg_frw_set": ([vp, C.c_double, C.c_double], C.c_int),
        "cg_frw_add_fluid": ([vp, C.c_double, C.c_double], C.c_int),
        "cg_frw_get": ([vp, dp], C.c_int),
        "cg_set_lambda": ([vp, C.c_double], C.c_int),
        "cg_set_gauge_params": ([vp, C.c_double, C.c_double, C.c_double], C.c_int),
        "cg_set_shift_params": ([vp, C.c_int, C.c_double, C.c_double, C.c_double], C.c_int),
        "cg_step_init": ([vp], C.c_int),
        "cg_rk_evolve": ([vp], C.c_int),
        "cg_k_finalize": ([vp, C.c_int], C.c_int),
        "cg_rk4_steps": ([vp, C.c_int], C.c_int),
        "cg_sync": ([vp], C.c_int),
        "cg_kernel_info": ([vp, ip, ip, ip, ip, ip, ip], C.c_int),
        "cg_set_timing": ([vp, C.c_int], C.c_int),
        "cg_set_dma": ([vp, C.c_int], C.c_int),
        "cg_set_fused_accumulator": ([vp, C.c_int], C.c_int),
        "cg_ics_static_dust": ([vp, dp, C.c_size_t, C.c_double, C.c_double, C.c_int, dp, C.POINTER(C.c_int)], C.c_int),
        "cg_set_tile_order": ([vp, C.c_int], C.c_int),
        "cg_set_dust_background": ([vp, C.c_int], C.c_int),
        "cg_autotune": ([vp, C.c_int, C.c_char_p, C.c_size_t], C.c_int),
        "cg_constraints": ([vp, dp, dp, dp, dp, dp], C.c_int),
        "cg_constraints_noalloc": ([vp, dp, dp, dp, dp, dp], C.c_int),
        "cg_bardeen": ([vp, C.c_double, C.c_int, C.c_double, dp], C.c_int),
        "cg_constraint_strip": ([vp, C.c_int, C.c_long, C.c_long, dp, dp, C.c_size_t], C.c_int),
        "cg_field_strip": ([vp, C.c_char_p, C.c_int, C.c_long, C.c_long, dp, C.c_size_t], C.c_int),
        "cg_field_stats": ([vp, C.c_char_p, dp], C.c_int),
        "cg_power_spectrum": ([vp, C.c_char_p, dp, C.c_int], C.c_int),
        "cg_stage_time_ms": ([vp, dp, ip], C.c_int),
        "cg_extra_field_data": ([vp, dp], C.c_int),
        "cg_batch_create": ([C.POINTER(_Params), C.c_int, C.POINTER(vp)], C.c_int),
        "cg_batch_destroy": ([vp], C.c_int),
        "cg_batch_last_error": ([vp], C.c_char_p),
        "cg_batch_size": ([vp], C.c_int),
        "cg_batch_member": ([vp, C.c_int, C.POINTER(vp)], C.c_int),
        "cg_batch_step": ([vp, C.c_int], C.c_int),
        "cg_batch_set_graph": ([vp, C.c_int], C.c_int),
        "cg_batch_set_dma": ([vp, C.c_int], C.c_int),
        "cg_batch_step_time_ms": ([vp, dp], C.c_int),
        "cg_batch_constraints": ([vp, dp, dp, dp, dp, dp], C.c_int),
        "cg_batch_autotune": ([vp, C.c_int, C.c_char_p, C.c_size_t], C.c_int),
        "cg_batch_kernel_info": ([vp, ip, ip, ip, ip, ip], C.c_int),
        "cg_num_nans": ([vp, C.c_char_p, C.POINTER(C.c_longlong)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes, f.restype = args, res
    _lib = L
    return L


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class BSSN:
    """GPU-resident BSSN state with the reference's method names."""

    def __init__(self, params: GpuParams, _borrowed=None):
        self.L = load_library()
        self.params = params
        self.shape = params.shape()
        self.npts = int(np.prod(self.shape))
        self._owner = None
        if _borrowed is not None:  # a member of a Batch: the batch owns the handle
            self._owner, self.h = _borrowed
            return
        h = C.c_void_p()
        rc = self.L.cg_create(C.byref(params.to_c()), C.byref(h))
        self.h = h
        if rc:
            msg = self.L.cg_last_error(h).decode() if h.value else "invalid parameters"
            if h.value:
                self.L.cg_destroy(h)
            self.h = None
            raise CosmoGpuError(rc, msg)
        self._chk(self.L.cg_set_gauge_params(self.h, params.gd_c, params.dw_mu_l, params.dw_p))
        self._chk(self.L.cg_set_shift_params(self.h, params.gauge_wave_dir, params.gd_eta, params.dw_mu_s,
                                             params.k_driver_coeff))

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_owner", None) is None:
            self.L.cg_destroy(self.h)
        self.h = None

    def _chk(self, rc):
        if rc:
            raise CosmoGpuError(rc, self.L.cg_last_error(self.h).decode())

    # --- fields map --------------------------------------------------
    def upload(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        if a.size != self.npts:
            raise ValueError(f"{name}: expected {self.npts} values, got {a.size}")
        self._chk(self.L.cg_field_upload(self.h, name.encode(), _dptr(a), a.size))

    def download(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self._chk(self.L.cg_field_download(self.h, name.encode(), _dptr(out), out.size))
        return out

    # --- FRW / Lambda --------------------------------------------------
    def frw_set(self, phi: float, K: float) -> None:
        self._chk(self.L.cg_frw_set(self.h, phi, K))

    def frw_add_fluid(self, rho: float, w: float) -> None:
        self._chk(self.L.cg_frw_add_fluid(self.h, rho, w))

    def frw_get(self):
        out = (C.c_double * 4)()
        self._chk(self.L.cg_frw_get(self.h, out))
        return tuple(out)

    def set_lambda(self, lam: float) -> None:
        self._chk(self.L.cg_set_lambda(self.h, lam))

    # --- RK4 methods (bssn.cc:219-339) -----------------------------------
    def stepInit(self):
        self._chk(self.L.cg_step_init(self.h))

    def RKEvolve(self):
        self._chk(self.L.cg_rk_evolve(self.h))

    def K1Finalize(self):
        self._chk(self.L.cg_k_finalize(self.h, 1))

    def K2Finalize(self):
        self._chk(self.L.cg_k_finalize(self.h, 2))

    def K3Finalize(self):
        self._chk(self.L.cg_k_finalize(self.h, 3))

    def K4Finalize(self):
        self._chk(self.L.cg_k_finalize(self.h, 4))

    def step(self, nsteps: int = 1):
        """nsteps whole RK4 steps (stepInit + 4 fused stages each)."""
        self._chk(self.L.cg_rk4_steps(self.h, nsteps))

    def setConstraintCalcs(self, noalloc: bool = False):
        """BSSN::setConstraintCalcs: dict of the H, M, G, A, S statistic arrays.
        noalloc=True: the same statistics through cg_constraints_noalloc
        (no device allocation; per-point arrays are not kept)."""
        H, M = (C.c_double * 8)(), (C.c_double * 8)()
        G, A, S = (C.c_double * 7)(), (C.c_double * 7)(), (C.c_double * 7)()
        fn = self.L.cg_constraints_noalloc if noalloc else self.L.cg_constraints
        self._chk(fn(self.h, H, M, G, A, S))
        return {"H": list(H), "M": list(M), "G": list(G), "A": list(A), "S": list(S)}

    def bardeen(self, t: float = 0.0, use_ml_scale_factor: bool = True, omega_L: float = 0.0) -> np.ndarray:
        """Bardeen::setPotentials at simulation time t; returns the 17 SVT
        violations (Bardeen::getSVTViolations).  The potentials are then
        readable as "Bardeen_Phi", "Bardeen_Psi", ... through download()."""
        out = np.zeros(17)
        self._chk(self.L.cg_bardeen(self.h, float(t), int(bool(use_ml_scale_factor)), float(omega_L), _dptr(out)))
        return out

    def set1DConstraintOutput(self, axis: int = 1, n1: int = 0, n2: int = 0):
        """BSSN::set1DConstraintOutput: (H, M_axis) at the points (x, n1, n2),
        x < N_axis (the running index is i for every axis, as in the reference)."""
        n = self.params.shape()[axis - 1]
        H, M = np.zeros(n), np.zeros(n)
        self._chk(self.L.cg_constraint_strip(self.h, int(axis), int(n1), int(n2), _dptr(H), _dptr(M), n))
        return H, M

    def strip(self, name: str, axis: int = 1, n1: int = 0, n2: int = 0) -> np.ndarray:
        """The values io_dump_strip writes for a field-map array along `axis`."""
        n = self.params.shape()[axis - 1]
        out = np.zeros(n)
        self._chk(self.L.cg_field_strip(self.h, name.encode(), int(axis), int(n1), int(n2), _dptr(out), n))
        return out

    def field_stats(self, name: str):
        """conformal_average, conformal_standard_deviation of a field and the
        volume_average of DIFFphi_a (io_bssn_dump_statistics' DETAILS)."""
        out = (C.c_double * 3)()
        self._chk(self.L.cg_field_stats(self.h, name.encode(), out))
        return tuple(out)

    def power_spectrum(self, name: str) -> np.ndarray:
        """Fourier::powerDump of a field-map array: |F_k|^2 averaged in bins of floor(|k|)."""
        nx, ny, nz = self.params.shape()
        nbins = int(np.sqrt(float(nx * nx + ny * ny + nz * nz)) / 2.0) + 1
        out = np.zeros(nbins)
        self._chk(self.L.cg_power_spectrum(self.h, name.encode(), _dptr(out), nbins))
        return out

    def extra_field_data(self):
        """BSSN::setExtraFieldData results: K_min, K_avg, avg_vol, rho_avg."""
        out = (C.c_double * 4)()
        self._chk(self.L.cg_extra_field_data(self.h, out))
        return dict(zip(("K_min", "K_avg", "avg_vol", "rho_avg"), tuple(out)))

    def numNaNs(self, name: str) -> int:
        n = C.c_longlong()
        self._chk(self.L.cg_num_nans(self.h, name.encode(), C.byref(n)))
        return n.value

    def sync(self):
        self._chk(self.L.cg_sync(self.h))

    # --- introspection / timing -----------------------------------------
    def kernel_info(self):
        v = [C.c_int() for _ in range(6)]
        self._chk(self.L.cg_kernel_info(self.h, *[C.byref(x) for x in v]))
        vg, sc, lds, ti, tj, tk = (x.value for x in v)
        return {"vgprs": vg, "scratch_bytes": sc, "lds_bytes": lds, "tile": (ti, tj, tk)}

    def autotune(self, reps: int = 3):
        """Time every compiled stage-kernel variant for this grid, keep the fastest."""
        import json
        buf = C.create_string_buffer(16384)
        self._chk(self.L.cg_autotune(self.h, reps, buf, len(buf)))
        return json.loads(buf.value.decode() or "[]")

    def set_timing(self, on: bool):
        self._chk(self.L.cg_set_timing(self.h, int(on)))

    def ics_static_dust(self, modes, rms: float, tol: float = 1e-15, max_iter: int = 200):
        """cg_ics_static_dust: static-dust initial data from the half-spectrum
        `modes` ([nx][ny][nz/2+1] complex), computed on the device.  Returns
        (shift c, iterations)."""
        m = np.ascontiguousarray(modes, dtype=np.complex128)
        flat = m.view(np.float64).reshape(-1)
        c, it = C.c_double(), C.c_int()
        self._chk(self.L.cg_ics_static_dust(self.h, _dptr(flat), flat.size, float(rms), float(tol), int(max_iter),
                                            C.byref(c), C.byref(it)))
        return c.value, it.value

    def set_dma(self, on: bool):
        """LDS-DMA halo staging where it applies (default on); off = register staging."""
        self._chk(self.L.cg_set_dma(self.h, int(on)))

    def set_dust_background(self, on: bool):
        """Dust fluid as perturbations of the FRW fluids (cg_set_dust_background)."""
        self._chk(self.L.cg_set_dust_background(self.h, int(on)))

    def set_fused_accumulator(self, keep: bool):
        """True: fused steps keep the reference's _f accumulator (17 register
        sweeps per step); False (default): the accumulator-free form (13)."""
        self._chk(self.L.cg_set_fused_accumulator(self.h, 1 if keep else 0))

    def set_tile_order(self, order: int):
        """Tile traversal order of the stage kernel (0-5; the autotuner sets it)."""
        self._chk(self.L.cg_set_tile_order(self.h, int(order)))

    def stage_time_ms(self):
        avg, n = C.c_double(), C.c_int()
        self._chk(self.L.cg_stage_time_ms(self.h, C.byref(avg), C.byref(n)))
        return avg.value, n.value


class Batch:
    """Batched ensemble (cg_batch_*): independent grids advanced by one
    stage-kernel launch per RK4 stage, replayed from a HIP graph.
    ``members[r]`` is a ``BSSN`` view of realisation r (uploads, downloads,
    FRW set-up, constraints...)."""

    def __init__(self, params_list):
        self.L = load_library()
        n = len(params_list)
        arr = (_Params * n)(*[p.to_c() for p in params_list])
        h = C.c_void_p()
        rc = self.L.cg_batch_create(arr, n, C.byref(h))
        self.h = h
        if rc:
            msg = self.L.cg_batch_last_error(h).decode() if h.value else "invalid parameters"
            if h.value:
                self.L.cg_batch_destroy(h)
            self.h = None
            raise CosmoGpuError(rc, msg)
        self.members = []
        for r, p in enumerate(params_list):
            mh = C.c_void_p()
            self._chk(self.L.cg_batch_member(self.h, r, C.byref(mh)))
            self.members.append(BSSN(p, _borrowed=(self, mh)))

    def __del__(self):
        if getattr(self, "h", None):
            for m in getattr(self, "members", []):
                m.h = None
            self.L.cg_batch_destroy(self.h)
            self.h = None

    def __len__(self):
        return len(self.members)

    def _chk(self, rc):
        if rc:
            raise CosmoGpuError(rc, self.L.cg_batch_last_error(self.h).decode())

    def step(self, nsteps: int = 1):
        self._chk(self.L.cg_batch_step(self.h, nsteps))

    def set_graph(self, on: bool):
        self._chk(self.L.cg_batch_set_graph(self.h, int(on)))

    def set_dma(self, on: bool):
        self._chk(self.L.cg_batch_set_dma(self.h, int(on)))

    def step_time_ms(self) -> float:
        ms = C.c_double()
        self._chk(self.L.cg_batch_step_time_ms(self.h, C.byref(ms)))
        return ms.value

    def setConstraintCalcs(self):
        """setConstraintCalcs of every realisation: dict of [n, 8|7] arrays."""
        n = len(self.members)
        H, M = np.zeros((n, 8)), np.zeros((n, 8))
        G, A, S = np.zeros((n, 7)), np.zeros((n, 7)), np.zeros((n, 7))
        self._chk(self.L.cg_batch_constraints(self.h, _dptr(H), _dptr(M), _dptr(G), _dptr(A), _dptr(S)))
        return {"H": H, "M": M, "G": G, "A": A, "S": S}

    def autotune(self, reps: int = 3):
        import json
        buf = C.create_string_buffer(16384)
        self._chk(self.L.cg_batch_autotune(self.h, reps, buf, len(buf)))
        return json.loads(buf.value.decode() or "[]")

    def kernel_info(self):
        v = [C.c_int() for _ in range(5)]
        self._chk(self.L.cg_batch_kernel_info(self.h, *[C.byref(x) for x in v]))
        vg, sc, ti, tj, tk = (x.value for x in v)
        return {"vgprs": vg, "scratch_bytes": sc, "tile": (ti, tj, tk)}
