"""Case driver: the reference program (main.cuf + solve_finite_volumes,
finite_volumes_method.cuf:4-258) for one subdomain, with the step on the GPU.

    python -m cuteflow_amd <case_dir>

reads <case_dir>/donnees.f, the mesh (text layout, or CGNS when is_cgns =
1), the boundary table (built and written when elt_bound = 0, read
otherwise), the initial condition (file, dam line or plane), runs to ts,
prints the status block every freqaffich steps and writes the snapshots
(legacy VTK, element/node text, CGNS FlowSolutions in out_<meshfile>) and
the restart file like the reference.
"""
from __future__ import annotations

import os
import shutil
import sys
import time

import numpy as np

from . import io as cio
from .mesh import read_boundary_table, read_text_mesh, write_boundary_table
from .params import read_donnees
from .solver import Solver, initial_state


class SnapshotWriter:
    """global_save_sol_visu (pre_post_traitement.cuf:5-49): snapshot k is
    written when tc is within +-tolaffiche of k*ts/visu_snapshots, with
    tolaffiche = 1.5*dt after the first one."""

    def __init__(self, prm, mesh, zm, surf, outdir, cgns_out=None):
        self.prm, self.mesh, self.zm, self.surf, self.outdir = prm, mesh, zm, surf, outdir
        self.times = cio.snapshot_times(prm.ts, prm.visu_snapshots) if prm.solvisu > 0 else np.zeros(0)
        self.cpt = 0
        self.tolaffiche = prm.tolaffiche
        self.written = []
        # CGNS export (pre_post_traitement.cuf:27-31) into the mesh copy
        self.cgns = None
        if cgns_out is not None and prm.is_cgns == 1 and prm.solvisu in (1, 3):
            from .cgns import CgnsExporter
            self.cgns = CgnsExporter(cgns_out)

    @property
    def pending(self):
        return self.cpt < len(self.times)

    def next_time(self):
        return self.times[self.cpt] if self.pending else np.inf

    def check(self, tc, dt, get_state):
        if not self.pending:
            return
        if self.cpt > 0:
            self.tolaffiche = 1.5 * dt
        t = self.times[self.cpt]
        if t - self.tolaffiche <= tc <= t + self.tolaffiche:
            self.write(self.cpt, tc, get_state())
            self.cpt += 1
        elif tc > t + self.tolaffiche:
            print("Skipped save sol ! : The timestep is too large for the number of snapshots to save!",
                  self.cpt)
            self.cpt += 1

    def write(self, k, tc, U):
        p, m, d = self.prm, self.mesh, self.outdir
        if p.solvisu in (1, 2):
            cio.write_vtk(os.path.join(d, f"_solution-Paraview_{k}.vtk"), m.coords, m.conn, U, self.zm,
                          p.tolisec, p.sol_z_offset)
        if p.solvisu in (1, 4):
            nodes = cio.solution_on_nodes(m.conn, m.nnt, U, self.surf)
            cio.write_simple_nodes(os.path.join(d, f"solution_noeuds_simple_{k}.txt"), tc, nodes,
                                   m.coords[:, 2], p.sol_z_offset)
        if self.cgns is not None:
            if k == 0:
                self.cgns.init(self.times)
            # the Manning field is written when the per-cell n is in use
            self.cgns.export(k, U, self.zm, m.manning if p.is_override_manning == 0 else None, p.tolisec,
                             p.sol_z_offset)
        if p.solvisu in (1, 5):
            cio.write_simple_elements(os.path.join(d, f"solution_elements_simple_{k}.txt"), tc, U, self.zm,
                                      p.sol_z_offset)
        self.written.append((k, tc))


def run_case(case_dir: str, verbose: bool = True, max_steps: int | None = None):
    """Runs a case directory; returns a dict with the final state and times."""
    prm = read_donnees(os.path.join(case_dir, "donnees.f"))
    mesh_path = os.path.join(case_dir, prm.meshfile_path, prm.meshfile.strip())
    cgns_out = None
    if prm.is_cgns:  # lec_mail_cgns (maillage.cuf:398-778)
        from .cgns import read_cgns_mesh
        mesh = read_cgns_mesh(mesh_path, prm.inlet_name, prm.outlet_name, prm.nombre_entrees,
                              prm.nombre_sorties, prm.is_override_manning, verbose)
        if prm.solvisu in (1, 3):  # the solutions go into a copy of the mesh file (maillage.cuf:427-437)
            cgns_out = os.path.join(case_dir, "out_" + prm.meshfile.strip())
            shutil.copyfile(mesh_path, cgns_out)
    else:
        mesh = read_text_mesh(mesh_path, prm.nombre_entrees, prm.nombre_sorties)
    prm.nombre_entrees, prm.nombre_sorties = mesh.nombre_entrees, mesh.nombre_sorties
    file_bc = os.path.join(case_dir, "boundary_table_" + prm.meshfile.strip())
    if prm.elt_bound == 0:
        bnd = mesh.boundary_table(device=True)
        write_boundary_table(file_bc, bnd)
    else:
        bnd = read_boundary_table(file_bc, mesh.nelt)
    s = Solver(mesh, prm, boundary=bnd)
    geo = s.geometry()
    if verbose:
        print(" cotemin             = ", geo["cotemin"])
        for k in range(prm.nombre_entrees):
            print(" debitglob       ", k + 1, " =", prm.debitglob[k])
            print(" longueur entree ", k + 1, " =", s.long_entree[k])
    if prm.solinit == 1:
        _, U0 = cio.read_initial_condition(os.path.join(case_dir, prm.fich_sol_init.strip()), mesh.nelt)
    else:
        U0 = initial_state(prm, mesh, geo["zm"])
    s.set_discharge(True)  # debit_entree / debit_sortie of the status block
    s.set_state(U0)
    snaps = SnapshotWriter(prm, mesh, geo["zm"], geo["surf"], case_dir, cgns_out)
    snaps.check(0.0, s.dt, s.get_state)

    freq = max(1, prm.freqaffich)
    nt, dts, q_in, q_out, steady = 0, [], [], [], False
    t0 = time.perf_counter()
    if verbose:
        print(" Loop over time starts")
    status = 0
    while True:
        # run in batches up to the next print; step by step close to a snapshot
        dt = s.dt
        n = freq - (nt % freq)
        if snaps.pending:
            room = snaps.next_time() - 1.5 * dt - s.tc
            n = min(n, max(1, int(room / (4.0 * dt)) if room > 0 else 1))
        if max_steps is not None:
            n = min(n, max_steps - nt)
            if n <= 0:
                break
        hist, done, st = s.step(n, raise_on_error=False)
        dts.extend(hist[:done])
        qi, qo, steady = s.discharge()
        q_in.extend(qi)
        q_out.extend(qo)
        nt += done
        if done and snaps.pending:
            snaps.check(s.tc, hist[done - 1], s.get_state)
        if verbose and steady:  # check_steady_state (finite_volumes_method.cuf:380-383)
            print(" -----------------------------------------------")
            print(" LE REGIME PERMANENT EST ATTEINT")
            print(" -----------------------------------------------")
        if verbose and done and nt % freq == 0:  # print_status_in_console (:346-363)
            print(" -----------------------------------------------")
            print(" Iteration number ", nt, ", t = ", s.tc)
            if prm.nombre_entrees == 1:
                print(" debit_entree ", 1, " = ", qi[-1])
            print(" debit_entree_sum =", qi[-1], "M3/S")
            print(" debit_sortie     =", qo[-1], "M3/S")
            print(" -----------------------------------------------")
        if st != 0:
            print(" Erreur sur le pas de temps dt=", s.dt)
            status = st
            break
        if done < n:
            break
    wall = time.perf_counter() - t0
    U = s.get_state()
    if verbose:
        print(" Loop on time ends at iteration : ", nt)
    if prm.solrestart == 1:
        cio.write_restart(os.path.join(case_dir, "solution_elements_restart.txt"), s.tc, U)
    if verbose:
        print(" DUREE DU CALCUL :", wall, " Secondes ")
    out = {"U": U, "tc": s.tc, "steps": nt, "dt_hist": np.array(dts), "status": status, "wall_s": wall,
           "snapshots": snaps.written, "zm": geo["zm"], "q_in": np.array(q_in), "q_out": np.array(q_out),
           "steady": steady}
    s.close()
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m cuteflow_amd <case_dir>", file=sys.stderr)
        return 2
    r = run_case(argv[0])
    return 0 if r["status"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
