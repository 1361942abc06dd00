"""Builds libcosmo_gpu.so in-tree with hipcc for gfx950.

    python -m cosmograph_amd.build

Every HIP translation unit is compiled with packed-FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcosmo_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("COSMO_GPU_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
SOURCES = ["stage_inst_f32.hip", "stage_inst_f32all.hip", "stage_inst_f32fix.hip", "stage_inst_f64.hip", "stage_inst_misc.hip", "stage_inst_dust.hip", "stage_inst_dustfix.hip", "cons_inst.hip",
           "cosmo_gpu.hip", "batch.hip", "ics.hip", "gauge.hip", "bardeen.hip", "fixmetric.hip"]


def _compile(src: str, build_dir: str, verbose: bool) -> str:
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return obj


def build(verbose: bool = True) -> str:
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir, verbose), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lhipfft"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    build_driver(verbose)
    return OUT


DRIVER = os.path.join(HERE, "cosmo_gpu")


def build_driver(verbose: bool = True) -> str:
    """The C++ host classes and the `cosmo_gpu` driver (plain host code:
    g++, linked against libcosmo_gpu.so next to it)."""
    host = os.path.join(HERE, "host")
    tmp = DRIVER + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(host, "cosmo_host.cc"), os.path.join(host, "cosmo_main.cc"),
           "-L", HERE, "-lcosmo_gpu", "-lz", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
