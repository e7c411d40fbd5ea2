"""Host (CPU) build of the step's cell body: compile + ctypes wrapper, shared
by the CPU derivative checks and the GPU tests that compare against it."""
import ctypes as C
import os
import subprocess

import numpy as np

from cuteflow_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
_DP = C.POINTER(C.c_double)


def build(outdir):
    """Compile host_adjoint.cpp with the library's host preprocessing into
    outdir and load it."""
    out = os.path.join(str(outdir), "libhost_adjoint.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                           "-Wno-unknown-pragmas", "-o", out,
                           os.path.join(HERE, "host_adjoint.cpp"),
                           os.path.join(ROOT, "cuteflow_amd", "csrc", "cf_prep.cpp")])
    lib = C.CDLL(out)
    lib.hs_create.restype = C.c_void_p
    lib.hs_create.argtypes = [C.POINTER(_lib.CfMesh), C.POINTER(_lib.CfParams)]
    lib.hs_destroy.argtypes = [C.c_void_p]
    lib.hs_set_ad_flags.argtypes = [C.c_void_p, C.c_int]
    lib.hs_set_manning.argtypes = [C.c_void_p, _DP]
    lib.hs_forward.restype = C.c_int
    lib.hs_forward.argtypes = [C.c_void_p, C.c_int, _DP, _DP, _DP]
    lib.hs_tlm.argtypes = [C.c_void_p, _DP, _DP, _DP, _DP]
    lib.hs_adjoint.restype = C.c_long
    lib.hs_adjoint.argtypes = [C.c_void_p, _DP, _DP, _DP]
    lib.hs_face_dot.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _DP, _DP, _DP,
                                C.c_double, C.c_double, C.c_double, C.c_double, _DP]
    lib.hs_muscl_face_dot.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _DP, _DP, _DP,
                                      C.c_double, C.c_double, C.c_double, C.c_double, _DP]
    return lib


def _p(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(_DP)


class Host:
    """Host build of the step on one mesh (caller's cell order)."""

    def __init__(self, lib, mesh, params):
        self.lib = lib
        self.bt = mesh.boundary_table()
        le = mesh.inlet_lengths(self.bt, params.nombre_entrees)
        self.cfp = params.to_cf(le, ts=1e300)
        self.conn = np.ascontiguousarray(mesh.conn.reshape(-1), np.int32)
        self.coords = np.ascontiguousarray(mesh.coords.reshape(-1), np.float64)
        self.mann = np.ascontiguousarray(mesh.manning, np.float64)
        cm = _lib.CfMesh()
        cm.nelt, cm.nnt = mesh.nelt, mesh.nnt
        cm.conn = _lib.iptr(self.conn)
        cm.coords = _lib.dptr(self.coords)
        cm.boundary = _lib.iptr(self.bt)
        cm.manning = _lib.dptr(self.mann)
        self.h = lib.hs_create(C.byref(cm), C.byref(self.cfp))
        assert self.h
        self.n = mesh.nelt

    def set_ad_flags(self, flags):
        self.lib.hs_set_ad_flags(self.h, int(flags))

    def set_manning(self, n):
        n = np.ascontiguousarray(n, np.float64)
        self.lib.hs_set_manning(self.h, _p(n))

    def forward(self, U0, nsteps):
        U0 = np.ascontiguousarray(U0, np.float64)
        hist = np.zeros(nsteps + 1)
        U = np.zeros((self.n, 3))
        done = self.lib.hs_forward(self.h, nsteps, _p(U0), hist.ctypes.data_as(_DP), U.ctypes.data_as(_DP))
        return U, hist[: done + 1], done

    def tlm(self, dn, dU0):
        dU = np.zeros((self.n, 3))
        U = np.zeros((self.n, 3))
        dn = np.ascontiguousarray(dn, np.float64)
        dU0 = np.ascontiguousarray(dU0, np.float64)
        self.lib.hs_tlm(self.h, _p(dn), _p(dU0), dU.ctypes.data_as(_DP), U.ctypes.data_as(_DP))
        return dU, U

    def adjoint(self, lam):
        lam = np.ascontiguousarray(lam, np.float64)
        gn = np.zeros(self.n)
        gU = np.zeros((self.n, 3))
        bad = self.lib.hs_adjoint(self.h, _p(lam), gn.ctypes.data_as(_DP), gU.ctypes.data_as(_DP))
        return gn, gU, bad

    def close(self):
        self.lib.hs_destroy(self.h)
