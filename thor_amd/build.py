"""Builds libthor_amd.so in-tree: host C++ with g++, HIP kernels with hipcc for gfx950.

Usage: python -m thor_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libthor_amd.so")
# Parity build: identical sources with the portable elementary functions shared
# with the CPU oracle (bit-for-bit comparison of the restated arithmetic).
LIB_PM = os.path.join(HERE, "libthor_amd_pm.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HOST_SRC = ["host.cpp", "grid.cpp"]
# THOR's file formats (.thr, HDF5 output and restart) and the CLI driver:
# built against the HDF5 C library of the image, beside the core library.
IO_LIB = os.path.join(HERE, "libthor_io.so")
ESP_EXE = os.path.join(HERE, "thor_esp")
HDF5 = os.environ.get("HDF5_ROOT", "/opt/conda")
HIP_SRC = ["dyn.hip", "mlp.hip", "grid_gpu.hip"]

CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
# The FP64 path must round every product and sum as written (no FMA
# contraction); packed FP32 instructions are disabled for all HIP code.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _link(objs, out):
    tmp = out + ".tmp"
    _run(["g++", "-shared", "-fopenmp", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64",
          f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, out)


def build(verbose: bool = True) -> str:
    """Compiles both variants of the library (production and parity build);
    the translation units compile in parallel."""
    os.makedirs(OBJ, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    jobs, host, objs = [], [], {LIB: [], LIB_PM: []}
    for src in HOST_SRC:
        o = os.path.join(OBJ, src + ".o")
        jobs.append(["g++", *CXXFLAGS, "-I", os.path.join(ROCM, "include"), "-c", os.path.join(CSRC, src), "-o", o])
        host.append(o)
    for lib, extra, tag in ((LIB, [], ""), (LIB_PM, ["-DTHOR_PORTABLE_MATH"], "_pm")):
        objs[lib] = list(host)
        for src in HIP_SRC:
            o = os.path.join(OBJ, src + tag + ".o")
            jobs.append([hipcc, *HIPFLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", o])
            objs[lib].append(o)
    with ThreadPoolExecutor(max_workers=max(1, min(len(jobs), os.cpu_count() or 1, 8))) as ex:
        list(ex.map(_run, jobs))
    for lib in (LIB, LIB_PM):
        _link(objs[lib], lib)
    build_io()
    return LIB


def build_io() -> None:
    """libthor_io.so (HDF5 and .thr) and the thor_esp driver, linked against
    libthor_amd.so through an rpath relative to themselves."""
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        print(f"HDF5 not found under {HDF5}: libthor_io.so and thor_esp are not built", flush=True)
        return
    rp = "-Wl,-rpath,$ORIGIN"
    # the system C++ runtime ahead of the HDF5 installation's own (older) one,
    # which ROCm's libraries cannot use
    sysdir = "/usr/lib/x86_64-linux-gnu"
    _run(["g++", *CXXFLAGS, "-shared", "-I", os.path.join(HDF5, "include"), os.path.join(CSRC, "esp_io.cpp"),
          "-o", IO_LIB + ".tmp", f"-L{HERE}", "-l:libthor_amd.so", f"-L{HDF5}/lib", "-lhdf5", rp,
          f"-Wl,-rpath,{sysdir}", f"-Wl,-rpath,{HDF5}/lib"])
    os.replace(IO_LIB + ".tmp", IO_LIB)
    _run(["g++", *CXXFLAGS, "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"),
          os.path.join(CSRC, "esp_main.cpp"), "-o", ESP_EXE + ".tmp", f"-L{HERE}", "-l:libthor_io.so",
          "-l:libthor_amd.so", f"-L{ROCM}/lib", "-lamdhip64", rp, f"-Wl,-rpath,{ROCM}/lib",
          f"-Wl,-rpath,{sysdir}", f"-Wl,-rpath,{HDF5}/lib"])
    os.replace(ESP_EXE + ".tmp", ESP_EXE)


if __name__ == "__main__":
    build()
    sys.exit(0)
