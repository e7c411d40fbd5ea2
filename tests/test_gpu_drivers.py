"""Case drivers on the GPU: the Fortran host (amdflang, iso_c_binding) and the
Python driver run the same case directory (donnees.f, text mesh, boundary
table, dam-line initial condition) and agree with the oracle; the Python
driver also runs a reference case directory as is (CGNS mesh, solvisu=3)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from cuteflow_amd import io as cio
from cuteflow_amd import mesh as M
from cuteflow_amd.build import FORTRAN_EXE
from cuteflow_amd.driver import run_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DONNEES = """&DONNEES_NAMELIST
  GP=9.81,
  is_override_manning=0, override_manning=0.0,
  is_cgns=0, meshfile_path='./', meshfile='reach.txt', elt_bound=0,
  nombre_entrees=1, nombre_sorties=1,
  eqlin_barrage=1, x1eqbar=150.0, x2eqbar=150.0001, y1eqbar=0.0, y2eqbar=200.0,
  H_AMONT={h_amont}, U_AMONT=0, V_AMONT=0, H_AVAL={h_aval}, U_AVAL=0, V_AVAL=0,
  plan=0, solinit=0, fich_sol_init='',
  inlet='{inlet}', inlet_name="in", debitglob=120.0,
  outlet="fixedheight", outlet_name="out", H_sortie={h_out},
  IFLUX=3, ilimiteur=1, iupwind=1, tolisec=1.0E-03, friction=1, fricimplic=1,
  TS={ts}, CFL=0.9, is_dt_constant=0, constant_dt=0.0, tol_reg_perm=0.0,
  freqaffich=50, sol_z_offset=0., solrestart=1, solvisu=5, visu_snapshots=2,
  sortie_finale_bluekenue=0 /
"""


def _make_case(tmp_path, ts=20.0, inlet="inflow"):
    r = M.river_mesh(nx=40, ny=10, length=400.0, width=200.0, channel_width=80.0, bar_len=100.0,
                     field_corr=50.0, seed=11)
    M.write_text_mesh(str(tmp_path / "reach.txt"), r.mesh)
    z = r.mesh.coords[:, 2]
    (tmp_path / "donnees.f").write_text(DONNEES.format(h_amont=float(z.max()) + 0.5, h_aval=float(z.max()) - 1.0,
                                                       h_out=r.h_out, ts=ts, inlet=inlet))
    return r.mesh


def _oracle(mesh, case_dir):
    from cuteflow_amd.params import read_donnees
    from cuteflow_amd.solver import initial_state
    p = read_donnees(os.path.join(case_dir, "donnees.f"))
    prep = O.Prepared(mesh)
    U0 = initial_state(p, mesh, prep.zm)
    return p, prep, O.run(prep, p, U0, 100000, qout=True)


def test_python_driver_matches_oracle(tmp_path):
    mesh = _make_case(tmp_path)
    out = run_case(str(tmp_path), verbose=False)
    p, prep, ref = _oracle(mesh, str(tmp_path))
    assert out["status"] == 0 and ref["status"] == 0
    assert out["steps"] == ref["steps"] and out["tc"] == ref["tc"]
    assert np.array_equal(out["U"], ref["U"])
    # per-step discharges of the status block
    assert np.array_equal(out["q_in"], ref["qin"]) and np.array_equal(out["q_out"], ref["qout"])
    assert not out["steady"]
    # boundary table file written like the reference
    bnd = M.read_boundary_table(str(tmp_path / "boundary_table_reach.txt"))
    assert np.array_equal(bnd, prep.boundary_fortran())
    # restart file: tc then (4f16.6) rows
    lines = open(tmp_path / "solution_elements_restart.txt").read().splitlines()
    assert float(lines[0]) == out["tc"] and len(lines) == mesh.nelt + 1
    assert lines[1] == cio._fmt(ref["U"][0])
    # snapshots at 0, ts/2, ts
    assert [k for k, _ in out["snapshots"]] == [0, 1, 2]


def test_fortran_driver_matches_python(tmp_path):
    if not os.path.exists(FORTRAN_EXE):
        pytest.skip("Fortran driver not built")
    mesh = _make_case(tmp_path)
    proc = subprocess.run([FORTRAN_EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "FIN DE LA SIMULATION" in proc.stdout
    fort_restart = open(tmp_path / "solution_elements_restart.txt").read().splitlines()
    fort_bt = open(tmp_path / "boundary_table_reach.txt").read()
    p, prep, ref = _oracle(mesh, str(tmp_path))
    # same step count and state (to the 6 decimals of the restart format)
    assert f"Loop on time ends at iteration :" in proc.stdout
    nt = int(proc.stdout.split("Loop on time ends at iteration :")[1].split()[0])
    assert nt == ref["steps"]
    assert abs(float(fort_restart[0]) - ref["tc"]) <= 1e-12 * ref["tc"]
    assert fort_restart[1:] == [cio._fmt(row) for row in ref["U"]]
    assert M.read_boundary_table(str(tmp_path / "boundary_table_reach.txt")).shape == (2, 3, mesh.nelt)
    assert fort_bt.splitlines()[0].split() == [str(v) for v in prep.boundary_fortran()[:, :, 0].reshape(-1)]
    # the status block prints each freqaffich-th step's discharges (list-directed: ~16 digits)
    printed = [float(line.split("=")[1].split()[0]) for line in proc.stdout.splitlines() if "debit_sortie" in line]
    iters = [int(line.split("Iteration number")[1].split(",")[0]) for line in proc.stdout.splitlines()
             if "Iteration number" in line]
    assert len(printed) == len(iters) > 0
    for it, q in zip(iters, printed):
        assert abs(q - ref["qout"][it - 1]) <= 1e-15 * max(abs(ref["qout"][it - 1]), 1e-300), (it, q)


def test_drivers_phys_inflow_match_oracle(tmp_path):
    """inlet='phys_inflow' (the inlet discharge shared over the inlet faces in
    proportion to the depth, flux_riemann.cuf:127-137) through both hosts:
    the Python driver's states, step count, tc and per-step discharges equal
    the oracle's bit for bit, the Fortran driver's restart rows and printed
    discharges equal the oracle's to the digits they print."""
    dirs = []
    for who in ("python", "fortran"):
        d = tmp_path / who
        d.mkdir()
        mesh = _make_case(d, inlet="phys_inflow")
        dirs.append(d)
    p, prep, ref = _oracle(mesh, str(dirs[0]))
    assert p.inlet == "phys_inflow" and ref["status"] == 0 and np.all(ref["qin"][1:] > 0)
    out = run_case(str(dirs[0]), verbose=False)
    assert out["status"] == 0 and out["steps"] == ref["steps"] and out["tc"] == ref["tc"]
    assert np.array_equal(out["U"], ref["U"])
    assert np.array_equal(out["q_in"], ref["qin"]) and np.array_equal(out["q_out"], ref["qout"])
    if not os.path.exists(FORTRAN_EXE):
        pytest.skip("Fortran driver not built")
    proc = subprocess.run([FORTRAN_EXE], cwd=str(dirs[1]), capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    nt = int(proc.stdout.split("Loop on time ends at iteration :")[1].split()[0])
    assert nt == ref["steps"]
    rows = open(dirs[1] / "solution_elements_restart.txt").read().splitlines()
    assert rows[1:] == [cio._fmt(row) for row in ref["U"]]
    lines = proc.stdout.splitlines()
    q_in = [float(line.split("=")[1].split()[0]) for line in lines if "debit_entree_sum" in line]
    iters = [int(line.split("Iteration number")[1].split(",")[0]) for line in lines if "Iteration number" in line]
    assert len(q_in) == len(iters) > 0
    for it, q in zip(iters, q_in):
        assert abs(q - ref["qin"][it - 1]) <= 1e-15 * abs(ref["qin"][it - 1]), (it, q)


def test_fortran_and_python_drivers_write_the_same_snapshots(tmp_path):
    """Both hosts on one case with every output on (solvisu=1, 3 snapshots):
    the snapshot scheduler (global_save_sol_visu) picks the same steps, and
    the VTK, node and element files are identical line for line (the
    list-directed time line compared as a number)."""
    if not os.path.exists(FORTRAN_EXE):
        pytest.skip("Fortran driver not built")
    dirs = []
    for who in ("python", "fortran"):
        d = tmp_path / who
        d.mkdir()
        _make_case(d, ts=12.0)
        src = (d / "donnees.f").read_text().replace("solvisu=5, visu_snapshots=2", "solvisu=1, visu_snapshots=3")
        (d / "donnees.f").write_text(src)
        dirs.append(d)
    out = run_case(str(dirs[0]), verbose=False)
    proc = subprocess.run([FORTRAN_EXE], cwd=str(dirs[1]), capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert [k for k, _ in out["snapshots"]] == [0, 1, 2, 3]
    for k in range(4):
        for name in (f"_solution-Paraview_{k}.vtk", f"solution_noeuds_simple_{k}.txt",
                     f"solution_elements_simple_{k}.txt"):
            a = (dirs[0] / name).read_text().splitlines()
            b = (dirs[1] / name).read_text().splitlines()
            if name.endswith(".txt"):
                assert float(a[0]) == float(b[0]), name
                a, b = a[1:], b[1:]
            assert a == b, name
    assert not list(dirs[1].glob("*_final*"))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_python_driver_reference_cgns_case(tmp_path):
    """The divergent canal case directory as the reference ran it (its
    donnees.f, the CGNS mesh file it names, solvisu=3), shortened to
    ts = 5 s with 4 snapshots: the state equals the oracle's on the text
    copy of the mesh, the boundary table file written is the reference's
    own, byte for byte, and out_<meshfile> receives the FlowSolutions."""
    from cuteflow_amd.build import CGNS_LIB_PATH
    if not os.path.exists(CGNS_LIB_PATH):
        pytest.skip("libcuteflow_cgns.so not built")
    src = open(os.path.join(GOLDEN, "divergent_canal.donnees.f"), encoding="latin-1").read()
    src = src.replace("TS=280.0", "TS=5.0").replace("visu_snapshots=100", "visu_snapshots=4")
    src = src.replace("freqaffich=1000", "freqaffich=50")
    (tmp_path / "donnees.f").write_text(src, encoding="latin-1")
    shutil.copyfile(os.path.join(GOLDEN, "divergent_canal.cgns"), tmp_path / "manning_Canal_divergent.cgns")
    print("cgns case: divergent canal, ts=5", flush=True)
    out = run_case(str(tmp_path), verbose=False)
    mesh = M.read_text_mesh(os.path.join(GOLDEN, "divergent_canal.mesh"))
    p, prep, ref = _oracle(mesh, str(tmp_path))
    assert out["status"] == 0 and ref["status"] == 0
    assert out["steps"] == ref["steps"] > 100 and out["tc"] == ref["tc"]
    assert np.array_equal(out["U"], ref["U"])
    bt = open(tmp_path / "boundary_table_manning_Canal_divergent.cgns").read()
    assert bt == open(os.path.join(GOLDEN, "divergent_canal.boundary_table")).read()
    written = [k for k, _ in out["snapshots"]]
    assert written == [0, 1, 2, 3, 4]
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not os.path.exists(h5dump):
        return
    listing = subprocess.run([h5dump, "-n", str(tmp_path / "out_manning_Canal_divergent.cgns")],
                             capture_output=True, text=True, check=True).stdout
    for k in written:
        assert f"/Base/Zone/Solution_{k}/Water Depth" in listing
    assert "Manning" not in listing.split("/Base/Zone/Solution_0", 1)[1]  # is_override_manning = 1
    k, tc = out["snapshots"][-1]
    if tc == out["tc"]:
        path = tmp_path / "out_manning_Canal_divergent.cgns"
        assert np.array_equal(_h5_field(path, f"/Base/Zone/Solution_{k}/Water Depth/ data"), out["U"][:, 0])


def _h5_field(path, ds):
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    dump = subprocess.run([h5dump, "-y", "-w", "0", "-m", "%.17g", "-d", ds, str(path)], capture_output=True,
                          text=True, check=True).stdout
    body = dump[dump.index("DATA {") + 6:]
    return np.array([float(v) for v in re.split(r"[,\s]+", body[: body.index("}")]) if v])


def test_fortran_driver_reference_cgns_case(tmp_path):
    """The Fortran host on the same reference case directory (CGNS mesh,
    solvisu=3): step count and restart file as the oracle, the reference's
    boundary table file byte for byte, the first FlowSolution in
    out_<meshfile> equal to the oracle's initial state, and every
    FlowSolution equal to the Python driver's."""
    from cuteflow_amd.build import CGNS_LIB_PATH
    from cuteflow_amd.solver import initial_state
    if not (os.path.exists(FORTRAN_EXE) and os.path.exists(CGNS_LIB_PATH)):
        pytest.skip("Fortran driver or libcuteflow_cgns.so not built")
    src = open(os.path.join(GOLDEN, "divergent_canal.donnees.f"), encoding="latin-1").read()
    src = src.replace("TS=280.0", "TS=5.0").replace("visu_snapshots=100", "visu_snapshots=4")
    src = src.replace("freqaffich=1000", "freqaffich=50").replace("solrestart=0", "solrestart=1")
    (tmp_path / "donnees.f").write_text(src, encoding="latin-1")
    shutil.copyfile(os.path.join(GOLDEN, "divergent_canal.cgns"), tmp_path / "manning_Canal_divergent.cgns")
    print("fortran cgns case: divergent canal, ts=5", flush=True)
    proc = subprocess.run([FORTRAN_EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "MANNING DETECTED IN MESH FILE BUT WILL NOT BE USED" in proc.stdout
    mesh = M.read_text_mesh(os.path.join(GOLDEN, "divergent_canal.mesh"))
    p, prep, ref = _oracle(mesh, str(tmp_path))
    nt = int(proc.stdout.split("Loop on time ends at iteration :")[1].split()[0])
    assert nt == ref["steps"] > 100
    lines = open(tmp_path / "solution_elements_restart.txt").read().splitlines()
    assert lines[1:] == [cio._fmt(row) for row in ref["U"]]
    bt = open(tmp_path / "boundary_table_manning_Canal_divergent.cgns").read()
    assert bt == open(os.path.join(GOLDEN, "divergent_canal.boundary_table")).read()
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not os.path.exists(h5dump):
        return
    out = tmp_path / "out_manning_Canal_divergent.cgns"
    U0 = initial_state(p, mesh, prep.zm)
    assert np.array_equal(_h5_field(out, "/Base/Zone/Solution_0/Water Depth/ data"), U0[:, 0])
    assert np.array_equal(_h5_field(out, "/Base/Zone/Solution_4/Bathymetry/ data"), prep.zm)
    # the Python driver on a copy of the case writes the same FlowSolutions
    py = tmp_path / "python"
    py.mkdir()
    shutil.copyfile(tmp_path / "donnees.f", py / "donnees.f")
    shutil.copyfile(os.path.join(GOLDEN, "divergent_canal.cgns"), py / "manning_Canal_divergent.cgns")
    pout = run_case(str(py), verbose=False)
    assert [k for k, _ in pout["snapshots"]] == [0, 1, 2, 3, 4]
    for k in range(5):
        for field in ("Water Depth", "VelocityX", "VelocityY"):
            ds = f"/Base/Zone/Solution_{k}/{field}/ data"
            assert np.array_equal(_h5_field(out, ds), _h5_field(py / "out_manning_Canal_divergent.cgns", ds)), ds
