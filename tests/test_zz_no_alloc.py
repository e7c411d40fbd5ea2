"""cf_build_boundary_device_noalloc (device pointers in and out, working
memory in a fixed per-GPU array, no device allocation) against
cf_build_boundary_device on the same meshes: equal tables, integer for
integer, on consecutive calls whose second mesh is smaller than the first
(the fixed array keeps the first call's contents)."""
import os

import numpy as np
import pytest

from cuteflow_amd import _lib
from cuteflow_amd import mesh as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAX_CELLS = 1 << 17


def _dev(a, torch):
    return torch.from_numpy(np.ascontiguousarray(a, np.int32)).cuda()


def _noalloc(m, torch, sentinel=-99):
    lib = _lib.load()
    conn = _dev(m.conn.reshape(-1), torch)
    lists = [_dev(x, torch) for x in (m.ndinput, m.numero_ndinput, m.ndoutput, m.numero_ndoutput, m.ndwall)]
    out = torch.full((6 * m.nelt,), sentinel, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    p = [x.data_ptr() if x.numel() else None for x in lists]
    st = lib.cf_build_boundary_device_noalloc(m.nelt, m.nnt, conn.data_ptr(), len(m.ndinput), p[0], p[1],
                                              len(m.ndoutput), p[2], p[3], len(m.ndwall), p[4], out.data_ptr(),
                                              stream)
    torch.cuda.synchronize()
    return st, out.cpu().numpy().reshape(2, 3, m.nelt)


def _meshes():
    big = M.river_mesh(nx=400, ny=80, length=2500.0, width=500.0, seed=7).mesh      # 64,000 cells
    small = M.read_text_mesh(os.path.join(GOLDEN, "river_piers.mesh"))              # 1,408 cells
    return big, small


@pytest.mark.gpu
def test_noalloc_boundary_equals_allocating_entry_point():
    import torch

    big, small = _meshes()
    assert big.nelt > 40 * small.nelt and big.nelt <= MAX_CELLS
    for m in (big, small):          # second call on a smaller mesh
        print(f"noalloc boundary: nelt={m.nelt} nnt={m.nnt}", flush=True)
        st, table = _noalloc(m, torch)
        assert st == _lib.CF_OK, _lib.load().cf_error_string(st)
        ref = m.boundary_table(device=True)
        assert np.array_equal(table, ref)
        assert np.array_equal(ref, m.boundary_table(device=False))
    # the smaller mesh's table is also the reference's own file
    assert np.array_equal(table, M.read_boundary_table(os.path.join(GOLDEN, "river_piers.boundary_table")))


@pytest.mark.gpu
def test_noalloc_boundary_refuses_large_and_invalid_meshes():
    import torch

    lib = _lib.load()
    conn = torch.ones(3, dtype=torch.int32, device="cuda")
    out = torch.zeros(6, dtype=torch.int32, device="cuda")
    # more cells than the fixed array holds: refused before any device work
    st = lib.cf_build_boundary_device_noalloc(MAX_CELLS + 1, 3, conn.data_ptr(), 0, None, None, 0, None, None,
                                              0, None, out.data_ptr(), None)
    assert st == _lib.CF_ERR_SIZE
    st = lib.cf_build_boundary_device_noalloc(1, MAX_CELLS + 1, conn.data_ptr(), 0, None, None, 0, None, None,
                                              0, None, out.data_ptr(), None)
    assert st == _lib.CF_ERR_SIZE
    # a node index outside 1..nnt: the mesh check of the allocating entry point
    bad = torch.tensor([1, 2, 7], dtype=torch.int32, device="cuda")
    st = lib.cf_build_boundary_device_noalloc(1, 3, bad.data_ptr(), 0, None, None, 0, None, None, 0, None,
                                              out.data_ptr(), None)
    assert st == _lib.CF_ERR_MESH
    # and the next call works again (the error flag is cleared per call)
    _, small = _meshes()
    st, table = _noalloc(small, torch)
    assert st == _lib.CF_OK and np.array_equal(table, small.boundary_table(device=True))
