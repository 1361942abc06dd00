"""cosmograph_amd -- MI355X-native BSSN evolution hot path of CosmoGRaPH.

The grid work runs in hand-written HIP kernels for gfx950 inside
``libcosmo_gpu.so`` (built in-tree by ``cosmograph_amd.build``), reached
through the C ABI declared in ``include/cosmo_gpu.h``.  This package loads
that library and fails loudly if it is missing: there is no CPU fallback.
"""
from .gpu import ERR_NAN, BSSN, Batch, CosmoGpuError, GpuParams, library_path, load_library  # noqa: F401
from .ics import static_dust_ics  # noqa: F401

__all__ = ["ERR_NAN", "BSSN", "Batch", "CosmoGpuError", "GpuParams", "library_path", "load_library", "static_dust_ics"]
