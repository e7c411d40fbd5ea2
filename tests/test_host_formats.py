"""Host-side formats and drivers without a GPU: namelist, text mesh,
boundary table file, initial-condition file and Fortran f16.6 output."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from cuteflow_amd import io as cio
from cuteflow_amd import mesh as M
from cuteflow_amd.params import f32, parse_namelist, read_donnees

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_namelist_reference_files():
    p = read_donnees(os.path.join(GOLDEN, "divergent_canal.donnees.f"))
    assert p.gp == 9.81 and p.is_override_manning == 1 and p.override_manning == 0.01
    assert p.iflux == 2 and p.ilimiteur == 3 and p.tolisec == 1.0e-5
    assert p.friction == 1 and p.fricimplic == 1 and p.ts == 280.0 and p.cfl == 0.9
    assert p.debitglob[0] == 0.35 and p.h_sortie[0] == 0.2
    assert p.inlet == "inflow" and p.outlet == "fixedheight"
    assert p.inlet_name[0] == "Inflow nodes" and p.outlet_name[0] == "Outflow nodes"
    assert p.plan == 1 and p.zpoint == 0.016 and p.xplan == 0.0
    assert p.solvisu == 3 and p.visu_snapshots == 100 and p.freqaffich == 1000
    # variables absent from the file keep the reference defaults
    assert p.tol == f32(1e-6) and p.tolaffiche == f32(1e-2)
    q = read_donnees(os.path.join(GOLDEN, "default.donnees.f"))
    assert q.iflux == 2 and q.debitglob[0] == 800.0 and q.h_sortie[0] == 29.0
    assert q.x1eqbar == 5.001 and q.xpoint == 274956.783790834
    assert q.meshfile == "mesh_7562_2.cgns" and q.is_cgns == 1
    r = read_donnees(os.path.join(GOLDEN, "river_piers.donnees.f"))
    assert r.iflux == 2 and r.tolisec == 1.0e-5


def test_namelist_syntax():
    nml = parse_namelist("""&donnees_namelist gp=9.81d0, debitglob=1.,2.5 h_sortie(3)=7, inlet='a/b',
        cfl=0.5 ! comment
        ts=3*1.0 /""")
    assert nml["gp"] == 9.81 and nml["debitglob"] == [1.0, 2.5] and nml["h_sortie"] == {2: 7}
    assert nml["inlet"] == "a/b" and nml["cfl"] == 0.5 and nml["ts"] == [1.0, 1.0, 1.0]


def test_text_mesh_roundtrip(tmp_path):
    m = M.read_text_mesh(os.path.join(GOLDEN, "river_piers.mesh"))
    assert (m.nnt, m.nelt, len(m.ndinput), len(m.ndoutput)) == (767, 1408, 21, 21)
    p = tmp_path / "m.txt"
    M.write_text_mesh(str(p), m)
    m2 = M.read_text_mesh(str(p))
    for f in ("coords", "conn", "manning", "ndinput", "ndoutput", "ndwall"):
        assert np.array_equal(getattr(m, f), getattr(m2, f)), f


@pytest.mark.parametrize("name", ["divergent_canal", "river_piers"])
def test_boundary_table_file_bytes(tmp_path, name):
    """The writer reproduces the reference's list-directed file byte for byte."""
    src = os.path.join(GOLDEN, f"{name}.boundary_table")
    bnd = M.read_boundary_table(src)
    out = tmp_path / "bt"
    M.write_boundary_table(str(out), bnd)
    assert open(out).read() == open(src).read()


def test_initial_condition_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    U = rng.normal(size=(50, 3))
    p = tmp_path / "ic.txt"
    cio.write_initial_condition(str(p), U, 7)
    tc0, V = cio.read_initial_condition(str(p), 50)
    assert tc0 == 7 and np.array_equal(U, V)


def test_f16_6_matches_fortran(tmp_path):
    """The f16.6 formatter gives the characters amdflang writes."""
    flang = "/opt/rocm/bin/amdflang"
    if not os.path.exists(flang):
        pytest.skip("amdflang not available")
    vals = [0.0, 1.0, -1.0, 1234.5678915, -0.0000004, 0.0000005, 2.5e-7, 123456789.123, -98765.4321, 1e-300,
            0.1234565, 7.0000005]
    src = tmp_path / "f.f90"
    src.write_text("program p\n  real(8) :: v(%d)\n  read(*,*) v\n  write(*,'(4f16.6)') v\nend program\n"
                   % len(vals))
    exe = tmp_path / "f"
    subprocess.run([flang, "-o", str(exe), str(src)], check=True, capture_output=True)
    got = subprocess.run([str(exe)], input=" ".join(repr(v) for v in vals), check=True,
                         capture_output=True, text=True).stdout.splitlines()
    ours = [cio._fmt(vals[i:i + 4]) for i in range(0, len(vals), 4)]
    assert got == ours


def test_snapshot_times():
    assert np.allclose(cio.snapshot_times(280.0, 100)[:3], [0.0, 2.8, 5.6])
    assert len(cio.snapshot_times(1.0, 0)) == 2
