"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, the smoke check and bench.py's CPU-baseline leg use this module;
it is the checker, never the product.  It does its own mesh preprocessing
(neighbours, boundary classes, geometry) with the oracle's C restatement of
the reference, so nothing here depends on the HIP library.

Inputs are duck-typed: a mesh needs coords (nnt,3), conn (nelt,3, 1-based),
manning (nelt), ndinput/numero_ndinput, ndoutput/numero_ndoutput, ndwall,
nombre_entrees; params needs the namelist variables used by the step.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAX_IO = 20
MAX_HYDRO = 64


class OrParams(C.Structure):
    _fields_ = [
        ("gp", C.c_double), ("tolisec", C.c_double), ("tol", C.c_double), ("cfl", C.c_double),
        ("ts", C.c_double), ("constant_dt", C.c_double), ("override_manning", C.c_double),
        ("tol_reg_perm", C.c_double),
        ("iflux", C.c_int), ("ilimiteur", C.c_int), ("friction", C.c_int), ("fricimplic", C.c_int),
        ("inlet", C.c_int), ("outlet", C.c_int), ("is_dt_constant", C.c_int),
        ("is_override_manning", C.c_int), ("nombre_entrees", C.c_int), ("pad0", C.c_int),
        ("debitglob", C.c_double * MAX_IO), ("h_sortie", C.c_double * MAX_IO),
        ("long_entree", C.c_double * MAX_IO),
        ("hydro_n", C.c_int * MAX_IO),
        ("hydro_t", C.c_double * (MAX_IO * MAX_HYDRO)),
        ("hydro_q", C.c_double * (MAX_IO * MAX_HYDRO)),
        ("surf_entree", C.c_double * MAX_IO),
    ]


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.or_pow.restype = C.c_double
        lib.or_pow.argtypes = [C.c_double, C.c_double]
        lib.or_cfl_dt.restype = C.c_double
        lib.or_run.restype = C.c_int
        lib.or_inflow_q.restype = C.c_double
        lib.or_inflow_q.argtypes = [C.POINTER(OrParams), C.c_int, C.c_double]
        lib.or_riemann_face.restype = None
        lib.or_riemann_face.argtypes = [C.c_int, C.c_int] + [C.c_double] * 17 + [_D]
        lib.or_limiter_value.restype = C.c_double
        lib.or_limiter_value.argtypes = [C.c_int, C.c_double, C.c_double]
        lib.or_slope_limiter_value.restype = C.c_double
        lib.or_slope_limiter_value.argtypes = [C.c_int, C.c_double, C.c_double]
        lib.or_muscl_face.restype = None
        lib.or_muscl_face.argtypes = [C.c_int] + [C.c_double] * 20 + [_D]
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def or_params(p, long_entree, ts=None) -> OrParams:
    o = OrParams()
    o.gp, o.tolisec, o.tol, o.cfl = p.gp, p.tolisec, p.tol, p.cfl
    o.ts = p.ts if ts is None else ts
    o.constant_dt, o.override_manning, o.tol_reg_perm = p.constant_dt, p.override_manning, p.tol_reg_perm
    o.iflux, o.ilimiteur, o.friction, o.fricimplic = p.iflux, p.ilimiteur, p.friction, p.fricimplic
    o.inlet = {"transm": 1, "inflow": 2, "phys_inflow": 3}[p.inlet.strip().lower()]
    o.outlet = {"fixedheight": 1, "transm": 2}[p.outlet.strip().lower()]
    o.is_dt_constant, o.is_override_manning = p.is_dt_constant, p.is_override_manning
    o.nombre_entrees = p.nombre_entrees
    for k in range(MAX_IO):
        o.debitglob[k] = float(p.debitglob[k])
        o.h_sortie[k] = float(p.h_sortie[k])
        o.long_entree[k] = float(long_entree[k]) if k < len(long_entree) else 0.0
    for inlet, (t, q) in getattr(p, "hydrographs", {}).items():
        k = int(inlet) - 1
        o.hydro_n[k] = len(t)
        for i, (ti, qi) in enumerate(zip(t, q)):
            o.hydro_t[k * MAX_HYDRO + i] = float(ti)
            o.hydro_q[k * MAX_HYDRO + i] = float(qi)
    return o


class Prepared:
    """Mesh products of the reference preprocessing, computed by the oracle."""

    def __init__(self, mesh, allpairs: bool | None = None):
        lib = load()
        n = int(mesh.conn.shape[0])
        self.nelt = n
        self.conn = np.ascontiguousarray(mesh.conn, np.int32).reshape(-1)
        coords = np.asarray(mesh.coords, np.float64)
        self.x = np.ascontiguousarray(coords[:, 0])
        self.y = np.ascontiguousarray(coords[:, 1])
        self.z = np.ascontiguousarray(coords[:, 2])
        self.manning = np.ascontiguousarray(mesh.manning, np.float64)
        self.b1 = np.zeros(3 * n, np.int32)
        self.b2 = np.zeros(3 * n, np.int32)
        if allpairs is None:
            allpairs = n <= 4000
        (lib.or_neighbours_allpairs if allpairs else lib.or_neighbours_sorted)(
            n, _i(self.conn), _i(self.b1), _i(self.b2))
        ndi = np.ascontiguousarray(mesh.ndinput, np.int32)
        ndo = np.ascontiguousarray(mesh.ndoutput, np.int32)
        ndw = np.ascontiguousarray(mesh.ndwall, np.int32)
        numi = np.ascontiguousarray(mesh.numero_ndinput, np.int32)
        numo = np.ascontiguousarray(mesh.numero_ndoutput, np.int32)
        lib.or_boundary_kind(n, _i(self.conn), _i(self.b1), _i(self.b2), len(ndi), _i(ndi), _i(numi),
                             len(ndo), _i(ndo), _i(numo), len(ndw), _i(ndw))
        # table as written to boundary_table_<mesh> (before walls)
        self.b1_table = self.b1.copy()
        lib.or_set_walls(n, _i(self.b1))
        self.upw = np.zeros(6 * n, np.int32)
        lib.or_upwind_cells(n, _i(self.b1), _i(self.upw))
        self.vx, self.vy = np.zeros(3 * n), np.zeros(3 * n)
        self.xc, self.yc, self.surf, self.zm, self.cot = (np.zeros(n) for _ in range(5))
        cmin = C.c_double()
        lib.or_geometry(n, _i(self.conn), _d(self.x), _d(self.y), _d(self.z), _d(self.vx), _d(self.vy),
                        _d(self.xc), _d(self.yc), _d(self.surf), _d(self.zm), _d(self.cot), C.byref(cmin))
        self.cotemin = cmin.value
        self.long_entree = np.zeros(MAX_IO)
        lib.or_long_entree(n, _i(self.conn), _d(self.x), _d(self.y), _i(self.b1), _i(self.b2),
                           MAX_IO, _d(self.long_entree))

    def boundary_fortran(self, walls: bool = False) -> np.ndarray:
        """boundary(nelt,3,2) as a (2,3,nelt) array."""
        b1 = self.b1 if walls else self.b1_table
        out = np.zeros((2, 3, self.nelt), np.int32)
        out[0] = b1.reshape(self.nelt, 3).T
        out[1] = self.b2.reshape(self.nelt, 3).T
        return out


def run(prep: Prepared, params, U0: np.ndarray, nsteps: int, ts: float | None = None,
        nthreads: int = 0, qout: bool = False):
    """The reference loop on the CPU: returns dict(U, dt_hist, steps, tc,
    status, steady[, qout, qin]) -- qout/qin the outflow/inflow discharge of
    each step's flux; steady whether the run stopped at steady state."""
    lib = load()
    p = or_params(params, prep.long_entree, ts=ts)
    U = np.ascontiguousarray(U0, np.float64).copy()
    dt_hist = np.zeros(nsteps + 1)
    qo = np.zeros(max(nsteps, 1)) if qout else None
    qi = np.zeros(max(nsteps, 1)) if qout else None
    tc = C.c_double()
    done = C.c_int()
    steady = C.c_int()
    st = lib.or_run(C.byref(p), prep.nelt, _i(prep.conn), _d(prep.vx), _d(prep.vy), _d(prep.xc),
                    _d(prep.yc), _d(prep.surf), _d(prep.zm), _d(prep.cot), _d(prep.manning),
                    _i(prep.b1), _i(prep.b2), _i(prep.upw), _d(U), int(nsteps), _d(dt_hist),
                    C.byref(tc), C.byref(done), _d(qo) if qout else None, _d(qi) if qout else None,
                    C.byref(steady), int(nthreads))
    out = {"U": U.reshape(-1, 3), "dt_hist": dt_hist[: done.value + 1], "steps": done.value,
           "tc": tc.value, "status": st, "steady": bool(steady.value)}
    if qout:
        out["qout"] = qo[: done.value]
        out["qin"] = qi[: done.value]
    return out


def riemann(iflux, left, right, n=(1.0, 0.0), gp=9.81, ilimiteur=1, upw=(0.0, 0.0, 0.0, 0.0),
            d=(1.0, 1.0, 1.0), dt=0.0):
    """Face flux (global frame) for face-frame states left=(h,un,ut), right=(h,un,ut)."""
    out = np.zeros(3)
    load().or_riemann_face(int(iflux), int(ilimiteur), gp, n[0], n[1], *left, *right, *upw, *d, dt, _d(out))
    return out


def limiter(ilimiteur, r, c):
    return load().or_limiter_value(int(ilimiteur), float(r), float(c))


def slope_limiter(ilimiteur, r1, r2):
    """simple_limiter: MUSCL slope from the upwind slope r1 and the slope r2 across the face."""
    return load().or_slope_limiter_value(int(ilimiteur), float(r1), float(r2))


def muscl(left, right, upw_left, upw_right, n=(1.0, 0.0), gp=9.81, ilimiteur=1, tolisec=1e-5,
          d=(1.0, 1.0, 1.0), dt=0.0):
    """MUSCL-Hancock face flux (global frame): face-frame states left/right=(h,un,ut), their
    upwind values upw_left/upw_right=(h,un,ut), d=(d_mrl, d_mkvol, d_mkvoisin)."""
    out = np.zeros(3)
    load().or_muscl_face(int(ilimiteur), gp, tolisec, n[0], n[1], *left, *right, *upw_left, *upw_right, *d,
                         dt, _d(out))
    return out


def inflow_q(params, inlet, tc):
    p = or_params(params, np.zeros(MAX_IO))
    return load().or_inflow_q(C.byref(p), int(inlet) - 1, float(tc))


def pow_(x: float, y: float) -> float:
    return load().or_pow(float(x), float(y))


def cfl_dt(prep: Prepared, params, U: np.ndarray) -> float:
    lib = load()
    p = or_params(params, prep.long_entree)
    U = np.ascontiguousarray(U, np.float64)
    return lib.or_cfl_dt(C.byref(p), prep.nelt, _d(prep.cot), _d(U))
