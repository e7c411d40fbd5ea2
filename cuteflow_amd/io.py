"""Result and initial-condition files in the reference's text formats.

Formats follow pre_post_traitement.cuf (simple_export :167-207,
paraview_export :407-496, save_sol_for_restart :498-512,
compute_sol_on_nodes :51-71) and read_initial_condition
(initial_condition.cuf:82-102).  Fortran's f16.6 edit descriptor is
reproduced by `_f16_6`.
"""
from __future__ import annotations

import numpy as np


def _f16_6(x: float) -> str:
    s = f"{x:16.6f}"
    if len(s) > 16:
        return "*" * 16
    if s.strip() in ("-0.000000",):
        s = f"{'-0.000000':>16}"
    return s


def _fmt(vals) -> str:
    return "".join(_f16_6(float(v)) for v in vals)


def _list_real(x: float) -> str:
    """A list-directed real as gfortran/flang print it (leading blank)."""
    return f" {float(x)!r}"


def read_initial_condition(path: str, nelt: int):
    """read_initial_condition: first line the integer tc_init, then nelt
    lines `h hu hv`; returns (tc_init, U (nelt, 3))."""
    with open(path, "r", encoding="latin-1") as fh:
        lines = fh.read().split("\n")
    tc_init = int(float(lines[0].split()[0].replace("d", "e").replace("D", "E")))
    vals = np.array(" ".join(lines[1:]).replace(",", " ").replace("d", "e").replace("D", "E").split(),
                    dtype=np.float64)
    if vals.size < 3 * nelt:
        raise ValueError(f"{path}: expected {nelt} cells, found {vals.size // 3}")
    return tc_init, vals[: 3 * nelt].reshape(nelt, 3)


def write_initial_condition(path: str, U: np.ndarray, tc_init: int = 0) -> None:
    """Full-precision file readable by read_initial_condition."""
    with open(path, "w") as fh:
        fh.write(f"{int(tc_init)}\n")
        for h, hu, hv in np.asarray(U, np.float64):
            fh.write(f"{float(h)!r} {float(hu)!r} {float(hv)!r}\n")


def write_restart(path: str, tc: float, U: np.ndarray) -> None:
    """save_sol_for_restart: tc, then `(4f16.6)` h hu hv per cell (lossy:
    6 decimals, as in the reference)."""
    with open(path, "w") as fh:
        fh.write(_list_real(tc) + "\n")
        for row in np.asarray(U, np.float64):
            fh.write(_fmt(row) + "\n")


def solution_on_nodes(conn: np.ndarray, nnt: int, U: np.ndarray, surf: np.ndarray) -> np.ndarray:
    """compute_sol_on_nodes: area-weighted average of the cells around each
    node, accumulated in cell order."""
    nodes = np.zeros((nnt, 3))
    area = np.zeros(nnt)
    for i in range(conn.shape[0]):
        for j in range(3):
            jj = conn[i, j] - 1
            nodes[jj] = nodes[jj] + U[i] * surf[i]
            area[jj] = area[jj] + surf[i]
    with np.errstate(invalid="ignore", divide="ignore"):
        return nodes / area[:, None]


def write_simple_elements(path: str, tc: float, U: np.ndarray, zm: np.ndarray, sol_z_offset: float = 0.0):
    """simple_export, element file: tc, then h, eta, hu, hv per cell."""
    with open(path, "w") as fh:
        fh.write(_list_real(tc) + "\n")
        for i in range(U.shape[0]):
            fh.write(_fmt((U[i, 0], U[i, 0] + zm[i] - sol_z_offset, U[i, 1], U[i, 2])) + "\n")


def write_simple_nodes(path: str, tc: float, nodes: np.ndarray, z: np.ndarray, sol_z_offset: float = 0.0):
    """simple_export, node file: tc, then h, eta, hu, hv per node."""
    with open(path, "w") as fh:
        fh.write(_list_real(tc) + "\n")
        for n in range(nodes.shape[0]):
            fh.write(_fmt((nodes[n, 0], nodes[n, 0] + z[n] - sol_z_offset, nodes[n, 1], nodes[n, 2])) + "\n")


def write_vtk(path: str, coords: np.ndarray, conn: np.ndarray, U: np.ndarray, zm: np.ndarray,
              tolisec: float, sol_z_offset: float = 0.0) -> None:
    """paraview_export: legacy ASCII VTK with eta, h, b, velocity and flux."""
    nnt, nelt = coords.shape[0], conn.shape[0]
    out = ["# vtk DataFile Version 1.0", "2D Unstructured Grid", "ASCII", "",
           "DATASET UNSTRUCTURED_GRID", f"POINTS{nnt:10d} float"]
    out += [_fmt((coords[n, 0], coords[n, 1], 0.0)) for n in range(nnt)]
    out += ["", f"CELLS{nelt:10d}{4 * nelt:10d}"]
    out += [f"3{conn[i, 0] - 1:10d}{conn[i, 1] - 1:10d}{conn[i, 2] - 1:10d}" for i in range(nelt)]
    out += ["", f"CELL_TYPES{nelt:10d}"] + ["5"] * nelt
    out += ["", f"CELL_DATA{nelt:10d}", "SCALARS eta float", "LOOKUP_TABLE default"]
    out += [_f16_6(U[i, 0] + zm[i] - sol_z_offset) for i in range(nelt)]
    out += ["", "SCALARS h float", "LOOKUP_TABLE default"] + [_f16_6(U[i, 0]) for i in range(nelt)]
    out += ["", "SCALARS b float", "LOOKUP_TABLE default"] + [_f16_6(zm[i] - sol_z_offset) for i in range(nelt)]
    out += ["", "VECTORS velocity float"]
    for i in range(nelt):
        if U[i, 0] > tolisec:
            out.append(_fmt((U[i, 1] / U[i, 0], U[i, 2] / U[i, 0], 0.0)))
        else:
            out.append(_fmt((0.0, 0.0, 0.0)))
    out += ["", "VECTORS flux float"]
    for i in range(nelt):
        out.append(_fmt((U[i, 1], U[i, 2], 0.0)) if U[i, 0] > tolisec else _fmt((0.0, 0.0, 0.0)))
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def snapshot_times(ts: float, visu_snapshots: int) -> np.ndarray:
    """tsolvisu(k) = k*ts/visu_snapshots, k = 0..visu_snapshots (main.cuf:113-122)."""
    n = max(int(visu_snapshots), 1)
    return np.array([k * ts / float(n) for k in range(n + 1)])
