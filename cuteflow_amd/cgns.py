"""CGNS mesh and solution files (libcuteflow_cgns.so, HDF5 layout).

read_cgns_mesh   lec_mail_cgns (maillage.cuf:398-778) on one rank
write_cgns_mesh  the converter's layout (cute_to_cgns.cpp:118-160) plus the
                 per-cell "Manning" user data (gen_manning.py:71-74)
CgnsExporter     cgns_export_init / cgns_export (pre_post_traitement.cuf:209-405)
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .mesh import Mesh


def _names(names, n: int, width: int = 100) -> bytes:
    """n fixed-width, blank-padded name fields (Fortran CHARACTER(width))."""
    out = b""
    for k in range(n):
        s = (names[k] if k < len(names) else "") or ""
        b = s.encode()[:width]
        out += b + b" " * (width - len(b))
    return out or b" "


def cgns_info(path: str) -> dict:
    """Sizes and contents of a mesh file, as cf_cgns_mesh_info reports them."""
    lib = _lib.load_cgns()
    info = _lib.CfCgnsInfo()
    _lib.check_cgns(lib.cf_cgns_mesh_info(str(path).encode(), info), "cf_cgns_mesh_info")
    return {n: getattr(info, n) for n, _ in _lib.CfCgnsInfo._fields_}


def read_cgns_mesh(path: str, inlet_name=(), outlet_name=(), nombre_entrees: int = 1, nombre_sorties: int = 1,
                   is_override_manning: int = 0, verbose: bool = False) -> Mesh:
    """Mesh of a CGNS file, numbered like the reference reader: inflow
    (outflow) nodes get the 1-based position of their BC's name among
    inlet_name[:nombre_entrees] (outlet_name[:nombre_sorties]), else 1.
    The per-cell Manning n comes from the file's "Manning" user data unless
    is_override_manning == 1 (zeros when not read)."""
    lib = _lib.load_cgns()
    info = cgns_info(path)
    nnt, nelt, ndi, ndo = info["nnt"], info["nelt"], info["ndi"], info["ndo"]
    coords = np.zeros((nnt, 3), np.float64)
    conn = np.zeros((nelt, 3), np.int32)
    mann = np.zeros(nelt, np.float64)
    ndin, numin = np.zeros(max(ndi, 1), np.int32), np.zeros(max(ndi, 1), np.int32)
    ndout, numout = np.zeros(max(ndo, 1), np.int32), np.zeros(max(ndo, 1), np.int32)
    width = 100
    st = lib.cf_cgns_read_mesh(str(path).encode(), nombre_entrees, _names(inlet_name, nombre_entrees, width),
                               nombre_sorties, _names(outlet_name, nombre_sorties, width), width,
                               int(is_override_manning), int(verbose), _lib.dptr(coords), _lib.iptr(conn),
                               _lib.dptr(mann), _lib.iptr(ndin), _lib.iptr(numin), _lib.iptr(ndout),
                               _lib.iptr(numout))
    _lib.check_cgns(st, "cf_cgns_read_mesh")
    return Mesh(coords=coords, conn=conn, manning=mann, ndinput=ndin[:ndi], numero_ndinput=numin[:ndi],
                ndoutput=ndout[:ndo], numero_ndoutput=numout[:ndo], manning_nd=np.zeros(nnt),
                nombre_entrees=nombre_entrees, nombre_sorties=nombre_sorties)


def write_cgns_mesh(path: str, m: Mesh, manning: bool = True) -> None:
    """Write m as a CGNS mesh file (with its per-cell n when manning)."""
    lib = _lib.load_cgns()
    coords = np.ascontiguousarray(m.coords, np.float64)
    conn = np.ascontiguousarray(m.conn, np.int32)
    man = np.ascontiguousarray(m.manning, np.float64)
    st = lib.cf_cgns_write_mesh(str(path).encode(), m.nnt, m.nelt, _lib.dptr(coords), _lib.iptr(conn),
                                m.nombre_entrees, len(m.ndinput), _lib.iptr(m.ndinput),
                                _lib.iptr(m.numero_ndinput), m.nombre_sorties, len(m.ndoutput),
                                _lib.iptr(m.ndoutput), _lib.iptr(m.numero_ndoutput), len(m.ndwall),
                                _lib.iptr(m.ndwall), _lib.dptr(man) if manning else None)
    _lib.check_cgns(st, "cf_cgns_write_mesh")


class CgnsExporter:
    """FlowSolution snapshots written into a copy of the mesh file (the
    reference's out_<meshfile>): init() once with the snapshot times, then
    export(k, ...) per snapshot written."""

    def __init__(self, path: str):
        self.path = str(path)
        self.lib = _lib.load_cgns()

    def init(self, times) -> None:
        t = np.ascontiguousarray(times, np.float64)
        _lib.check_cgns(self.lib.cf_cgns_export_init(self.path.encode(), len(t), _lib.dptr(t)),
                        "cf_cgns_export_init")

    def export(self, k: int, U: np.ndarray, zm: np.ndarray, manning: np.ndarray | None, tolisec: float,
               sol_z_offset: float = 0.0) -> None:
        U = np.ascontiguousarray(U, np.float64)
        zm = np.ascontiguousarray(zm, np.float64)
        man = None if manning is None else np.ascontiguousarray(manning, np.float64)
        st = self.lib.cf_cgns_export(self.path.encode(), int(k), U.shape[0], _lib.dptr(U), _lib.dptr(zm),
                                     None if man is None else _lib.dptr(man), float(tolisec),
                                     float(sol_z_offset))
        _lib.check_cgns(st, "cf_cgns_export")
