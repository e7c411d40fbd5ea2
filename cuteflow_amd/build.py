"""Build recipes for the native parts of cuteflow_amd.

The HIP library is compiled for gfx950 with hipcc directly (no torch
extension machinery, no JIT cache): the shared object lands in-tree under
cuteflow_amd/lib so that it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(LIB_DIR, "libcuteflow_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction, so the kernels round every operation
# exactly like the Fortran-order CPU evaluation they are checked against.
# Packed FP32 instructions are disabled for all HIP code of this project.
HIP_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC",
    "-shared",
    "-Wall",
]

SOURCES = [os.path.join(CSRC, "cuteflow_hip.hip"), os.path.join(CSRC, "cf_adjoint.hip"),
           os.path.join(CSRC, "cf_autotune.hip"), os.path.join(CSRC, "cf_prep.cpp"),
           os.path.join(CSRC, "cf_prep_gpu.hip")]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build_hip(verbose: bool = False) -> str:
    """Compile libcuteflow_hip.so for gfx950 (always recompiles)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


CGNS_SOURCE = os.path.join(CSRC, "cf_cgns.cpp")
CGNS_LIB_PATH = os.path.join(LIB_DIR, "libcuteflow_cgns.so")

FORTRAN_DIR = os.path.join(PKG_DIR, "fortran")
BIN_DIR = os.path.join(PKG_DIR, "bin")
FORTRAN_EXE = os.path.join(BIN_DIR, "cuteflow_hip")
AMDFLANG = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = False) -> str:
    """Compile the Fortran driver (amdflang) against libcuteflow_hip.so, and
    against libcuteflow_cgns.so (CGNS meshes, CF_WITH_CGNS) when that library
    has been built."""
    os.makedirs(BIN_DIR, exist_ok=True)
    obj = os.path.join(PKG_DIR, "_build", "fortran")
    os.makedirs(obj, exist_ok=True)
    with_cgns = os.path.exists(CGNS_LIB_PATH)
    mod_o = os.path.join(obj, "cuteflow_hip_mod.o")
    cgns_o = os.path.join(obj, "cuteflow_cgns_mod.o")
    main_o = os.path.join(obj, "cuteflow_main.o")
    steps = [[AMDFLANG, "-O2", "-module-dir", obj, "-c", os.path.join(FORTRAN_DIR, "cuteflow_hip_mod.f90"),
              "-o", mod_o]]
    objs = [mod_o]
    libs = ["-lcuteflow_hip"]
    if with_cgns:
        steps.append([AMDFLANG, "-O2", "-module-dir", obj, "-c", os.path.join(FORTRAN_DIR, "cuteflow_cgns_mod.f90"),
                      "-o", cgns_o])
        objs.append(cgns_o)
        libs.insert(0, "-lcuteflow_cgns")
    steps.append([AMDFLANG, "-O2", "-cpp", *(["-DCF_WITH_CGNS"] if with_cgns else []), "-module-dir", obj, "-I",
                  obj, "-c", os.path.join(FORTRAN_DIR, "cuteflow_main.f90"), "-o", main_o])
    steps.append([AMDFLANG, "-o", FORTRAN_EXE + ".tmp", *objs, main_o, "-L" + LIB_DIR, *libs,
                  "-Wl,-rpath,$ORIGIN/../lib"])
    for cmd in steps:
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
    os.replace(FORTRAN_EXE + ".tmp", FORTRAN_EXE)
    return FORTRAN_EXE


HDF5_DIR = os.environ.get("CF_HDF5_DIR", "/opt/conda")


def hdf5_available() -> bool:
    return os.path.exists(os.path.join(HDF5_DIR, "include", "hdf5.h")) and any(
        f.startswith("libhdf5.so") for f in os.listdir(os.path.join(HDF5_DIR, "lib")))


def build_cgns(verbose: bool = False) -> str | None:
    """Compile libcuteflow_cgns.so (host C++ over the HDF5 C API; the HDF5
    library is found at run time through an rpath to HDF5_DIR/lib).  Returns
    None, with a message, when no HDF5 installation is found."""
    if not hdf5_available():
        print(f"cuteflow_amd: no HDF5 under {HDF5_DIR}; the CGNS reader/writer is not built")
        return None
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = CGNS_LIB_PATH + ".tmp"
    inc, lib = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-I" + inc, "-o", tmp,
           CGNS_SOURCE, "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib, "-static-libstdc++", "-static-libgcc"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, CGNS_LIB_PATH)
    return CGNS_LIB_PATH


if __name__ == "__main__":
    print(build_hip(verbose=True))
    print(build_cgns(verbose=True))
    print(build_fortran(verbose=True))
