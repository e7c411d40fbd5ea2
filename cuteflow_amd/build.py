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
           os.path.join(CSRC, "cf_autotune.hip"), os.path.join(CSRC, "cf_prep.cpp")]


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


FORTRAN_DIR = os.path.join(PKG_DIR, "fortran")
BIN_DIR = os.path.join(PKG_DIR, "bin")
FORTRAN_EXE = os.path.join(BIN_DIR, "cuteflow_hip")
AMDFLANG = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = False) -> str:
    """Compile the Fortran driver (amdflang) against libcuteflow_hip.so."""
    os.makedirs(BIN_DIR, exist_ok=True)
    obj = os.path.join(PKG_DIR, "_build", "fortran")
    os.makedirs(obj, exist_ok=True)
    mod_o = os.path.join(obj, "cuteflow_hip_mod.o")
    main_o = os.path.join(obj, "cuteflow_main.o")
    steps = [
        [AMDFLANG, "-O2", "-module-dir", obj, "-c", os.path.join(FORTRAN_DIR, "cuteflow_hip_mod.f90"), "-o", mod_o],
        [AMDFLANG, "-O2", "-module-dir", obj, "-I", obj, "-c", os.path.join(FORTRAN_DIR, "cuteflow_main.f90"),
         "-o", main_o],
        [AMDFLANG, "-o", FORTRAN_EXE + ".tmp", mod_o, main_o, "-L" + LIB_DIR, "-lcuteflow_hip",
         "-Wl,-rpath,$ORIGIN/../lib"],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
    os.replace(FORTRAN_EXE + ".tmp", FORTRAN_EXE)
    return FORTRAN_EXE


if __name__ == "__main__":
    print(build_hip(verbose=True))
    print(build_fortran(verbose=True))
