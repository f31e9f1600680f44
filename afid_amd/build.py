"""Builds libafid_hip.so (the HIP engine behind include/afid.h) in-tree, and
the Fortran host afid_run on top of it.

hipcc compiles each translation unit for gfx950 and links one shared library
into afid_amd/lib/.  Packed-FP32 instructions are disabled for all device code.
g++ builds afid_amd/io/afid_io.cpp (continuation files, HDF5 C library) into
afid_amd/lib/libafid_io.so.  amdflang compiles afid_amd/fortran/
(iso_c_binding module + main program) into afid_amd/lib/afid_run, linked
against both libraries with an $ORIGIN rpath.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libafid_hip.so")
SOURCES = ["column_kernels.hip", "fft_kernels.hip", "misc_kernels.hip", "step_kernels.hip", "afid_ctx.hip"]
ARCH = os.environ.get("AFID_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def _hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    hipcc = _hipcc()

    def compile_one(src: str) -> str:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
        if verbose and r.stderr:
            print(r.stderr)
        return obj

    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    r = subprocess.run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    build_io()
    build_fortran_host()
    return LIB


IO_LIB = os.path.join(LIBDIR, "libafid_io.so")
HDF5_ROOT = os.environ.get("AFID_HDF5_ROOT", "/opt/conda")


def build_io() -> str:
    """libafid_io.so: continuation files over the C ABI and HDF5's C library."""
    tmp = IO_LIB + ".tmp"
    # libstdc++ is linked statically: the HDF5 directory on the rpath also holds
    # an older libstdc++ that must not be the one the HIP runtime gets.
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-static-libstdc++", "-static-libgcc",
           f"-I{HDF5_ROOT}/include",
           os.path.join(HERE, "io", "afid_io.cpp"), f"-L{LIBDIR}", "-lafid_hip", f"-L{HDF5_ROOT}/lib", "-lhdf5",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{HDF5_ROOT}/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for afid_io.cpp:\n{r.stderr[-4000:]}")
    os.replace(tmp, IO_LIB)
    return IO_LIB


FORTRAN_SOURCES = ["afid_hip.F90", "afid_main.F90"]
RUN = os.path.join(LIBDIR, "afid_run")


def build_fortran_host() -> str:
    fc = os.environ.get("AFID_FC", "/opt/rocm/bin/amdflang")
    fdir = os.path.join(HERE, "fortran")
    moddir = os.path.join(LIBDIR, "obj", "fmod")
    os.makedirs(moddir, exist_ok=True)
    tmp = RUN + ".tmp"
    cmd = [fc, "-cpp", "-O2", "-J", moddir, *[os.path.join(fdir, f) for f in FORTRAN_SOURCES],
           f"-L{LIBDIR}", "-lafid_io", "-lafid_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, RUN)
    return RUN


if __name__ == "__main__":
    print(build(verbose=True))
