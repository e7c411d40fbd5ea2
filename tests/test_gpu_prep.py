"""Mesh preprocessing on the GPU (cf_build_boundary_device, the device
preparation cf_create uses) against the host code, bit for bit, on the
reference's calibration meshes (whose boundary tables are the reference's own
files), synthetic meshes with several inlets and outlets, walls, a
non-manifold edge and duplicated boundary nodes; the error statuses agree."""
import ctypes as C
import os
import time

import numpy as np
import pytest

from cuteflow_amd import _lib
from cuteflow_amd import mesh as M
from cuteflow_amd.params import Params

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _meshes():
    out = {}
    for name in ("divergent_canal", "river_piers"):
        out[name] = M.read_text_mesh(os.path.join(GOLDEN, f"{name}.mesh"))
    r = M.river_mesh(nx=40, ny=12, length=400.0, width=120.0, channel_width=40.0, bar_len=100.0, seed=4)
    m = r.mesh
    m.numero_ndinput = (np.arange(len(m.ndinput)) % 3 + 1).astype(np.int32)
    m.numero_ndoutput = (np.arange(len(m.ndoutput)) // 4 % 2 + 1).astype(np.int32)
    m.nombre_entrees, m.nombre_sorties = 3, 2
    out["river_multi_io"] = m
    out["dam_walls"] = M.dam_break_mesh(30, jitter=0.3, seed=2)
    # shuffled cells, a duplicated inflow node with another number, some wall nodes
    q = M.river_mesh(nx=30, ny=8, seed=6).mesh
    q = q.permuted(np.random.default_rng(1).permutation(q.nelt))
    q.ndinput = np.concatenate([q.ndinput, q.ndinput[:2]])
    q.numero_ndinput = np.concatenate([np.ones(len(q.ndinput) - 2, np.int32), np.full(2, 2, np.int32)])
    q.nombre_entrees = 2
    q.ndwall = np.unique(np.concatenate([q.ndwall, q.conn[:5].reshape(-1)])).astype(np.int32)
    out["shuffled_dup_nodes"] = q
    # a non-manifold edge: a third triangle on an interior edge (the
    # neighbour taken is the highest cell id holding the reversed edge)
    d = M.dam_break_mesh(6, jitter=0.0, seed=0)
    a, b = d.conn[10, 0], d.conn[10, 1]
    extra = np.array([[b, a, 1], [a, b, 2]], np.int32)
    d.conn = np.concatenate([d.conn, extra])
    d.manning = np.zeros(d.nelt)
    out["non_manifold"] = d
    return out


MESHES = _meshes()


@pytest.mark.parametrize("name", list(MESHES))
def test_device_boundary_table_equals_host(name):
    m = MESHES[name]
    host = m.boundary_table(device=False)
    dev = m.boundary_table(device=True)
    assert np.array_equal(host, dev)
    if name in ("divergent_canal", "river_piers"):
        ref = M.read_boundary_table(os.path.join(GOLDEN, f"{name}.boundary_table"), m.nelt)
        assert np.array_equal(dev, ref)


def _prepare(m, bnd, params, tile, keep, device):
    lib = _lib.load()
    conn = np.ascontiguousarray(m.conn.reshape(-1), np.int32)
    coords = np.ascontiguousarray(m.coords.reshape(-1), np.float64)
    bnd = np.ascontiguousarray(bnd, np.int32)
    cm = _lib.CfMesh()
    cm.nelt, cm.nnt = m.nelt, m.nnt
    cm.conn, cm.coords, cm.boundary, cm.manning = _lib.iptr(conn), _lib.dptr(coords), _lib.iptr(bnd), None
    cp = params.to_cf(m.inlet_lengths(bnd, params.nombre_entrees))
    n = m.nelt
    nt = (n + tile - 1) // tile
    perm, nbr, cn = np.zeros(n, np.int32), np.zeros(4 * n, np.int32), np.zeros(3 * n, np.int32)
    off, adj = np.zeros(nt + 1, np.int32), np.zeros(3 * n + 1, np.int32)
    alw, io = np.zeros(nt, np.uint8), np.zeros(n, np.int32)
    nadj, nio = C.c_int(), C.c_int()
    st = lib.cf_debug_prepare(C.byref(cm), C.byref(cp), tile, int(keep), int(device), _lib.iptr(perm), _lib.iptr(nbr),
                              _lib.iptr(cn), _lib.iptr(off), _lib.iptr(adj), len(adj), C.byref(nadj),
                              alw.ctypes.data_as(C.c_void_p), _lib.iptr(io), C.byref(nio))
    if st:
        return st
    return perm, nbr, cn, off, adj[: nadj.value], alw, io[: nio.value]


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("tile,keep", [(256, 0), (64, 0), (128, 1)])
def test_device_prepare_equals_host(name, tile, keep):
    m = MESHES[name]
    bnd = m.boundary_table(device=False)
    p = Params(inlet="inflow", nombre_entrees=max(1, int(m.nombre_entrees)), nombre_sorties=1)
    h = _prepare(m, bnd, p, tile, keep, 0)
    d = _prepare(m, bnd, p, tile, keep, 1)
    assert not isinstance(h, int) and not isinstance(d, int), (h, d)
    for a, b, what in zip(h, d, ("perm", "nbr", "conn", "tile_off", "tile_adj", "tile_always", "io_cells")):
        assert np.array_equal(a, b), what
    if not keep:
        assert not np.array_equal(h[0], np.arange(m.nelt))  # a real reordering was compared


def test_device_prepare_errors_match_host():
    m = M.river_mesh(nx=20, ny=8, length=200.0, width=80.0, channel_width=40.0, bar_len=50.0, seed=8).mesh
    bnd = m.boundary_table(device=False)
    p = Params(inlet="inflow", nombre_entrees=1, nombre_sorties=1)
    # an unknown boundary class at a later cell, an inlet of zero length at an earlier one
    bad = bnd.copy()
    bad[0, 2, m.nelt - 3] = -5
    assert _prepare(m, bad, p, 64, 0, 0) == _prepare(m, bad, p, 64, 0, 1) == _lib.CF_ERR_MESH
    inlet_cells = np.nonzero((bnd[0] == -1).any(axis=0))[0]
    bad2 = bad.copy()
    bad2[1, :, inlet_cells[0]] = np.where(bad2[0, :, inlet_cells[0]] == -1, 2, bad2[1, :, inlet_cells[0]])
    assert inlet_cells[0] < m.nelt - 3
    # inlet 2 has no length: CF_ERR_ARG, met first in cell order
    assert _prepare(m, bad2, p, 64, 0, 0) == _prepare(m, bad2, p, 64, 0, 1) == _lib.CF_ERR_ARG
    # out-of-range neighbour id
    bad3 = bnd.copy()
    bad3[0, 1, 5] = m.nelt + 7
    assert _prepare(m, bad3, p, 64, 0, 0) == _prepare(m, bad3, p, 64, 0, 1) == _lib.CF_ERR_MESH
    # connectivity out of range
    lib = _lib.load()
    conn = np.ascontiguousarray(m.conn.reshape(-1), np.int32).copy()
    conn[7] = m.nnt + 1
    out = np.zeros(6 * m.nelt, np.int32)
    z = np.zeros(1, np.int32)
    for fn in (lib.cf_build_boundary, lib.cf_build_boundary_device):
        assert fn(m.nelt, m.nnt, _lib.iptr(conn), 0, _lib.iptr(z), None, 0, _lib.iptr(z), None, 0, _lib.iptr(z),
                  _lib.iptr(out)) == _lib.CF_ERR_MESH


def test_device_preprocessing_time_2m_cells():
    """The 2M-cell reach: boundary table and layout on the GPU against the
    host code (printed; the tables are equal)."""
    r = M.river_mesh(nx=3200, ny=320, length=20000.0, width=2000.0, seed=1)
    m = r.mesh
    t = time.perf_counter()
    host = m.boundary_table(device=False)
    t_host = time.perf_counter() - t
    m.boundary_table(device=True)  # first call: device set-up
    t = time.perf_counter()
    dev = m.boundary_table(device=True)
    t_dev = time.perf_counter() - t
    assert np.array_equal(host, dev)
    p = Params(inlet="inflow", nombre_entrees=1, nombre_sorties=1)
    t = time.perf_counter()
    h = _prepare(m, host, p, 64, 0, 0)
    t_ph = time.perf_counter() - t
    t = time.perf_counter()
    d = _prepare(m, host, p, 64, 0, 1)
    t_pd = time.perf_counter() - t
    for a, b in zip(h, d):
        assert np.array_equal(a, b)
    print(f"2M cells: boundary table host {t_host:.3f} s, device {t_dev:.3f} s; "
          f"layout host {t_ph:.3f} s, device {t_pd:.3f} s (incl. transfers)", flush=True)
