"""CGNS mesh reader and writers (libcuteflow_cgns.so, host only).

The reader is pinned to the reference's own data:
- the HDF5 originals of the two calibration meshes read into the same arrays
  as the text fixtures that h5dump extracted from them;
- the boundary table built from the CGNS-read meshes equals the reference's
  own boundary_table files.
The writer is checked node by node against the reference's mesh file (h5dump
of both, integer storage widths aside).  The FlowSolution export is checked
against values computed here from the definitions of
pre_post_traitement.cuf:344-366 (no reference output file is available:
the solution files in the reference tree are Git-LFS stubs)."""
import dataclasses
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from cuteflow_amd import _lib
from cuteflow_amd import mesh as M
from cuteflow_amd.build import CGNS_LIB_PATH, hdf5_available

pytestmark = pytest.mark.skipif(not (hdf5_available() and os.path.exists(CGNS_LIB_PATH)),
                                reason="libcuteflow_cgns.so not built (no HDF5)")

from cuteflow_amd import cgns  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
needs_h5dump = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
CASES = ["divergent_canal", "river_piers"]


def h5_values(path, ds, fmt="%.17g"):
    """Values of one dataset, decoded by h5dump (independent of this build)."""
    out = subprocess.run([H5DUMP, "-y", "-w", "0", "-m", fmt, "-d", ds, str(path)], capture_output=True,
                         text=True, check=True).stdout
    body = out[out.index("DATA {") + 6: out.rindex("}\n}")]
    body = body[: body.rindex("}")]
    return [v for v in re.split(r"[,\s]+", body) if v]


def h5_tree(path, rename):
    """Whole-file h5dump in creation order, with the file name replaced."""
    out = subprocess.run([H5DUMP, "-w", "0", "-y", "-q", "creation_order", str(path)], capture_output=True,
                         text=True, check=True).stdout
    return out.replace(str(path), rename)


@pytest.mark.parametrize("case", CASES)
def test_reader_matches_text_fixture(case):
    m = cgns.read_cgns_mesh(os.path.join(GOLDEN, f"{case}.cgns"), ["Inflow nodes"], ["Outflow nodes"])
    t = M.read_text_mesh(os.path.join(GOLDEN, f"{case}.mesh"))
    for f in ("coords", "conn", "manning", "ndinput", "ndoutput"):
        assert np.array_equal(getattr(m, f), getattr(t, f)), f
    assert (m.numero_ndinput == 1).all() and (m.numero_ndoutput == 1).all()
    assert len(m.ndwall) == 0


@pytest.mark.parametrize("case", CASES)
def test_reader_boundary_table_is_the_reference_one(case):
    """tab_bound_tri + boundary_kind on the CGNS-read mesh reproduce the
    boundary_table file the reference wrote for this mesh."""
    m = cgns.read_cgns_mesh(os.path.join(GOLDEN, f"{case}.cgns"), ["Inflow nodes"], ["Outflow nodes"])
    ref = M.read_boundary_table(os.path.join(GOLDEN, f"{case}.boundary_table"), m.nelt)
    assert np.array_equal(m.boundary_table(), ref)


def test_info():
    i = cgns.cgns_info(os.path.join(GOLDEN, "river_piers.cgns"))
    assert i == {"nnt": 767, "nelt": 1408, "ndi": 21, "ndo": 21, "nzones": 1, "zone": 1, "has_manning": 1,
                 "partitioned": 0}
    i = cgns.cgns_info(os.path.join(GOLDEN, "mesh_6072.cgns"))
    assert (i["nnt"], i["nelt"], i["ndi"], i["ndo"], i["has_manning"]) == (3237, 6072, 0, 0, 0)


@needs_h5dump
def test_reader_against_h5dump_no_bcs():
    """A mesh with no ZoneBC and no Manning data: every array as h5dump
    decodes it, no inflow/outflow nodes, n = 0 (nothing read)."""
    path = os.path.join(GOLDEN, "mesh_6072.cgns")
    m = cgns.read_cgns_mesh(path)
    for d, c in enumerate("XYZ"):
        v = np.array([float(x) for x in h5_values(path, f"/Base/Zone/GridCoordinates/Coordinate{c}/ data")])
        assert np.array_equal(m.coords[:, d], v)
    conn = np.array([int(x) for x in h5_values(path, "/Base/Zone/Elements/ElementConnectivity/ data", "%d")])
    assert np.array_equal(m.conn.reshape(-1), conn)
    assert len(m.ndinput) == 0 and len(m.ndoutput) == 0 and not m.manning.any()
    # with no inflow or outflow node no face is an inlet (-1) or outlet (-2)
    bnd = m.boundary_table()
    assert not np.isin(bnd[0], [-1, -2]).any() and (bnd[0] > 0).any()


def test_reader_bc_names(capfd):
    path = os.path.join(GOLDEN, "divergent_canal.cgns")
    # the BC's position among the names; the last match wins; trailing blanks do not count
    m = cgns.read_cgns_mesh(path, ["other", "Inflow nodes  "], ["Outflow nodes", "x", "Outflow nodes"],
                            nombre_entrees=2, nombre_sorties=3, verbose=True)
    assert (m.numero_ndinput == 2).all() and (m.numero_ndoutput == 3).all()
    assert "not found" not in capfd.readouterr().out
    # unmatched: 1, with the reference's message
    m = cgns.read_cgns_mesh(path, ["a", "b"], ["c"], nombre_entrees=2, nombre_sorties=1, verbose=True)
    assert (m.numero_ndinput == 1).all() and (m.numero_ndoutput == 1).all()
    out = capfd.readouterr().out
    assert "Inlet name : Inflow nodes, not found in donnees.f file ! Setting it to 1" in out
    assert "Outlet name : Outflow nodes, not found in donnees.f file ! Setting it to 1" in out
    # leading blanks of a name field are significant (Fortran comparison)
    m = cgns.read_cgns_mesh(path, [" Inflow nodes"], ["Outflow nodes"])
    assert (m.numero_ndinput == 1).all()


def test_reader_manning_override(capfd):
    path = os.path.join(GOLDEN, "divergent_canal.cgns")
    m = cgns.read_cgns_mesh(path, is_override_manning=1, verbose=True)
    assert not m.manning.any()
    assert "MANNING DETECTED IN MESH FILE BUT WILL NOT BE USED" in capfd.readouterr().out
    m = cgns.read_cgns_mesh(path, is_override_manning=0, verbose=True)
    assert m.manning.min() > 0
    assert "MANNING DETECTED IN MESH FILE WILL BE USED" in capfd.readouterr().out


def test_reader_errors(tmp_path):
    with pytest.raises(_lib.CuteflowError) as e:
        cgns.read_cgns_mesh(str(tmp_path / "missing.cgns"))
    assert e.value.status == _lib.CF_ERR_IO and "cannot open" in str(e.value)
    # the reference's boundary_table_*.cgns files are text, not HDF5
    with pytest.raises(_lib.CuteflowError) as e:
        cgns.cgns_info(os.path.join(GOLDEN, "divergent_canal.boundary_table"))
    assert e.value.status == _lib.CF_ERR_IO


@needs_h5dump
def test_writer_layout_is_the_reference_one(tmp_path):
    """Written from the arrays read out of the reference's file, the new
    file holds the same nodes, attributes, data and creation order; only the
    integer width of ElementRange/ElementConnectivity differs (I4 written,
    I8 in the reference file; CGNS readers take either)."""
    src = os.path.join(GOLDEN, "divergent_canal.cgns")
    m = cgns.read_cgns_mesh(src, ["Inflow nodes"], ["Outflow nodes"])
    out = tmp_path / "w.cgns"
    cgns.write_cgns_mesh(str(out), m)
    ref = h5_tree(src, "F").replace("H5T_STD_I64LE", "H5T_STD_I32LE").replace('"I8"', '"I4"')
    ours = h5_tree(out, "F")
    ours = re.sub(r"HDF5 Version [0-9.]+", "", ours)
    ref = re.sub(r"HDF5 Version [0-9.]+", "", ref)
    assert ours == ref


def test_writer_roundtrip_multi_inlets(tmp_path):
    """Two inlets and two outlets: one BC per number, named as the
    converter names them; read back with those names the numbers return."""
    r = M.river_mesh(nx=12, ny=8, length=120.0, width=40.0, channel_width=20.0, bar_len=30.0, seed=3)
    m = r.mesh
    assert len(m.ndinput) >= 4 and len(m.ndoutput) >= 4
    # inlet 1 / 2 on the two halves of the inflow node list, interleaved numbers on the outflow
    m.numero_ndinput = (np.arange(len(m.ndinput)) >= len(m.ndinput) // 2).astype(np.int32) + 1
    m.numero_ndoutput = (np.arange(len(m.ndoutput)) % 2 + 1).astype(np.int32)
    m.nombre_entrees = m.nombre_sorties = 2
    m.ndwall = np.arange(1, 6, dtype=np.int32)
    out = tmp_path / "multi.cgns"
    cgns.write_cgns_mesh(str(out), m)
    back = cgns.read_cgns_mesh(str(out), ["Inflow nodes 0", "Inflow nodes 1"],
                               ["Outflow nodes 0", "Outflow nodes 1"], 2, 2)
    assert np.array_equal(back.coords, m.coords) and np.array_equal(back.conn, m.conn)
    assert np.array_equal(back.manning, m.manning)
    for nd, num, nd0, num0 in ((back.ndinput, back.numero_ndinput, m.ndinput, m.numero_ndinput),
                               (back.ndoutput, back.numero_ndoutput, m.ndoutput, m.numero_ndoutput)):
        # BC order: all nodes of number 1, then number 2, each in the caller's order
        order = np.argsort(num0, kind="stable")
        assert np.array_equal(nd, nd0[order]) and np.array_equal(num, num0[order])
    assert len(back.ndwall) == 0  # BCWall is not read (ndwall = 0 in lec_mail_cgns)
    # the same inflow faces and numbers as the mesh it came from (the outflow
    # numbering was reordered, and walls are not read)
    b0, b1 = back.boundary_table(), dataclasses.replace(m, ndwall=np.zeros(0, np.int32)).boundary_table()
    assert np.array_equal(b0[0], b1[0])
    inflow = b0[0] == -1
    assert inflow.any() and np.array_equal(b0[1][inflow], b1[1][inflow])


def test_single_point_bc_is_ignored(tmp_path):
    """BCs of one point are skipped (n_bcnodes > 1, maillage.cuf:541, :577)."""
    r = M.river_mesh(nx=12, ny=8, length=120.0, width=40.0, channel_width=20.0, bar_len=30.0, seed=3)
    m = r.mesh
    m.numero_ndinput = np.ones(len(m.ndinput), np.int32)
    m.numero_ndinput[-1] = 2
    m.nombre_entrees = 2
    out = tmp_path / "one.cgns"
    cgns.write_cgns_mesh(str(out), m)
    assert cgns.cgns_info(str(out))["ndi"] == len(m.ndinput) - 1
    back = cgns.read_cgns_mesh(str(out), ["Inflow nodes 0", "Inflow nodes 1"], ["Outflow nodes"], 2, 1)
    assert np.array_equal(back.ndinput, m.ndinput[:-1]) and (back.numero_ndinput == 1).all()


@needs_h5dump
def test_flow_solution_export(tmp_path):
    path = tmp_path / "out_canal.cgns"
    shutil.copyfile(os.path.join(GOLDEN, "divergent_canal.cgns"), path)
    m = cgns.read_cgns_mesh(str(path))
    rng = np.random.default_rng(5)
    n = m.nelt
    U = np.stack([rng.uniform(0, 0.3, n), rng.normal(size=n), rng.normal(size=n)], 1)
    U[::7, 0] = 1e-6  # dry cells (h <= tolisec)
    zm = rng.normal(size=n)
    tol, off = 1e-5, 0.25
    ex = cgns.CgnsExporter(str(path))
    times = np.array([0.0, 140.0, 280.0])
    ex.init(times)
    ex.export(0, U, zm, m.manning, tol, off)
    ex.export(2, 2 * U, zm, None, tol, off)
    ex.export(2, U, zm, None, tol, off)  # rewriting a snapshot replaces it

    def field(k, name):
        return np.array([float(v) for v in h5_values(path, f"/Base/Zone/Solution_{k}/{name}/ data")])

    wet = U[:, 0] > tol
    for k in (0, 2):
        assert np.array_equal(field(k, "Water Depth"), U[:, 0])
        assert np.array_equal(field(k, "Bathymetry"), zm)
        assert np.array_equal(field(k, "Free surface elevation"), (U[:, 0] + zm) - off)
        assert np.array_equal(field(k, "MomentumX"), np.where(wet, U[:, 1], 0.0))
        assert np.array_equal(field(k, "MomentumY"), np.where(wet, U[:, 2], 0.0))
        with np.errstate(divide="ignore", invalid="ignore"):
            assert np.array_equal(field(k, "VelocityX"), np.where(wet, U[:, 1] / U[:, 0], 0.0))
            assert np.array_equal(field(k, "VelocityY"), np.where(wet, U[:, 2] / U[:, 0], 0.0))
        assert not field(k, "MomentumZ").any() and not field(k, "VelocityZ").any()
    assert np.array_equal(field(0, "Manning"), m.manning)
    listing = subprocess.run([H5DUMP, "-n", str(path)], capture_output=True, text=True, check=True).stdout
    assert "/Base/Zone/Solution_2/Manning" not in listing and "/Base/Zone/Solution_1" not in listing
    assert [float(v) for v in h5_values(path, "/Base/TimeIterValues/TimeValues/ data")] == list(times)
    assert h5_values(path, "/Base/TimeIterValues/ data", "%d") == ["3"]
    loc = bytes(int(v) for v in h5_values(path, "/Base/Zone/Solution_0/GridLocation/ data", "%d"))
    assert loc == b"CellCenter"
    ptr = bytes(int(v) for v in h5_values(path, "/Base/Zone/ZoneIterativeData/FlowSolutionPointers/ data", "%d"))
    assert ptr == b"".join(f"Solution_{k}".ljust(32).encode() for k in range(3))
    # the mesh is untouched
    m2 = cgns.read_cgns_mesh(str(path))
    assert np.array_equal(m2.coords, m.coords) and np.array_equal(m2.conn, m.conn)
    # wrong cell count
    with pytest.raises(_lib.CuteflowError):
        ex.export(1, U[:-1], zm[:-1], None, tol, off)
