"""Builds adga_amd/libadga_amd.so (HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python -m adga_amd.build   (or adga_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libadga_amd.so")
SOURCES = ["local.hip", "bse.hip", "lu.hip", "eom.hip", "kconv.hip", "onep.hip", "vq.hip", "modules.hip", "tangent.hip",
           "tail.hip", "lambda.hip", "inverse_noalloc.hip", "api.cpp", "api_modules.cpp", "api_tangent.cpp",
           "api_lambda.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         # packed FP32 instructions stay off (hipcc's host pass warns that it does not know the feature)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + CSRC, "-I" + os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return out


FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
MKL_DIR = os.environ.get("ADGA_LAPACK_DIR", "/opt/conda/lib")
FORTRAN = ["adga_c.f90", "lapack_module.f90", "adga_gpu.f90", "fortran_check.f90"]
FCHECK = os.path.join(HERE, "fortran", "adga_fortran_check")


def build_fortran() -> str | None:
    """Fortran host layer (ISO_C_BINDING modules) and its check program, linked to the library."""
    if not os.path.exists(FC):
        return None
    fdir = os.path.join(OBJ, "fortran")
    os.makedirs(fdir, exist_ok=True)
    objs = []
    for src in FORTRAN:
        out = os.path.join(fdir, os.path.splitext(src)[0] + ".o")
        r = subprocess.run([FC, "-O2", "-fPIC", "-module-dir", fdir, "-c", os.path.join(HERE, "fortran", src),
                            "-o", out], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {src}:\n{r.stderr[-4000:]}")
        objs.append(out)
    # LAPACK for lapack_module's host path above the GPU inverse's largest size (MKL of the image)
    r = subprocess.run([FC, "-o", FCHECK, *objs, "-L" + HERE, "-ladga_amd", "-Wl,-rpath,$ORIGIN/..",
                        "-L" + MKL_DIR, "-lmkl_rt", "-Wl,-rpath," + MKL_DIR],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran link failed:\n{r.stderr[-4000:]}")
    return FCHECK


H5_DIR = os.environ.get("ADGA_HDF5_DIR", "/opt/conda")
H5LIB = os.path.join(HERE, "libadga_h5.so")


def build_h5() -> str | None:
    """HDF5 readers and writers of the reference's file formats (host C++, libadga_h5.so)."""
    if not os.path.exists(os.path.join(H5_DIR, "include", "hdf5.h")):
        return None
    tmp = H5LIB + ".tmp"
    r = subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-I" + os.path.join(H5_DIR, "include"),
                        "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "h5io.cpp"), "-o", tmp,
                        "-L" + os.path.join(H5_DIR, "lib"), "-lhdf5", "-Wl,-rpath," + os.path.join(H5_DIR, "lib")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for h5io.cpp:\n{r.stderr[-4000:]}")
    os.replace(tmp, H5LIB)
    return H5LIB


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    fexe = build_fortran()
    h5 = build_h5()
    if verbose:
        print("built", LIB, "and", fexe or "(no amdflang: Fortran layer skipped)", "and",
              h5 or "(no HDF5 headers: HDF5 I/O skipped)")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
