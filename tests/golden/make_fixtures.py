#!/usr/bin/env python3
"""Extracts the reference-held fixtures of the two calibration cases into
small data files under tests/golden/ (run once, in a container where the
reference tree exists; the outputs are committed).

Sources (relative to the reference root), data only:
  Calibration_manning/docs_calibration/<case>/Froude/manning_<mesh>.cgns
      HDF5 CGNS mesh: coordinates, TRI_3 connectivity, Inflow/Outflow node
      PointLists, per-cell "Manning" array (read with the HDF5 h5dump tool)
  Calibration_manning/docs_calibration/<case>/Froude/boundary_table_manning_<mesh>.cgns
      the reference's own boundary(nelt,3,2) table (a text file despite its name)
  Calibration_manning/docs_calibration/<case>/Froude/outfile*.out
      GPU run logs: known answers (cotemin, tol, counts, inlet length, series)
  Calibration_manning/docs_calibration/<case>/Froude/donnees.f, Solution/Manning.txt,
      Solution/Entree_gen_manning.txt, src/cuteflow/donnees.f: input files
  Calibration_manning/Exemple2/base_files/mesh_6072.cgns: a CGNS mesh without BCs
      or Manning data (copied as is, like the two manning_<mesh>.cgns files)

Outputs: <case>.mesh (the text mesh layout of maillage.cuf:196-396),
<case>.cgns (the HDF5 original, copied), <case>.boundary_table (copied),
<case>.expected.json, <case>.donnees.f, <case>.Manning.txt,
<case>.Entree_gen_manning.txt, default.donnees.f, mesh_6072.cgns.
"""
import json
import os
import re
import shutil
import subprocess

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
H5DUMP = "/opt/conda/bin/h5dump"
CASES = {
    "divergent_canal": ("Calibration_manning/docs_calibration/Divergent_Canal", "Canal_divergent",
                        ["outfile_manning_min.out", "outfile_manning_max.out"]),
    "river_piers": ("Calibration_manning/docs_calibration/River_piers", "River_wth_piers", ["outfile.out"]),
}


def h5(path, ds, fmt="%.17g"):
    out = subprocess.run([H5DUMP, "-y", "-w", "0", "-m", fmt, "-d", ds, path], capture_output=True,
                         text=True, check=True, cwd="/tmp").stdout
    body = out[out.index("DATA {") + 6: out.rindex("}\n}")]
    body = body[: body.rindex("}")]
    return [v for v in re.split(r"[,\s]+", body) if v]


def parse_log(path):
    text = open(path, encoding="latin-1").read()
    exp = {}
    for key, pat in (("cotemin", r"cotemin\s*=\s*(\S+)"), ("tol", r"\btol\s*=\s*(\S+)"),
                     ("tolisec", r"tolisec\s*=\s*(\S+)"), ("nelt", r"nombre d'elements totaux =\s*(\d+)"),
                     ("debitglob_1", r"debitglob\s+1\s*=\s*(\S+)"),
                     ("long_entree_1", r"longueur entree\s+1\s*=\s*(\S+)"),
                     ("h_sortie_1", r"H_SORTIE\s+1\s*=\s*(\S+)"), ("iflux", r"iflux\s*=\s*(\d+)"),
                     ("loop_end_iteration", r"Loop on time ends at iteration :\s*(\d+)"),
                     ("duree_s", r"DUREE DU CALCUL :\s*(\S+)")):
        m = re.search(pat, text)
        if m:
            exp[key] = m.group(1)
    m = re.search(r"gridx, gridy.*\n\s*\d+ tab_bound_tri ok\s+(\d+)\s+(\d+)", text)
    if m:
        exp["nelt_nnt"] = [int(m.group(1)), int(m.group(2))]
    series = re.findall(r"Iteration number\s+(\d+)\s*, t =\s*(\S+)\s*\n\s*debit_entree\s+1\s*=\s*(\S+)\s*\n"
                        r"\s*debit_entree_sum\s*=\s*(\S+)\s*M3/S\s*\n\s*debit_sortie\s*=\s*(\S+)", text)
    exp["series"] = [[int(a), float(b), float(c), float(d), float(e)] for a, b, c, d, e in series]
    return exp


def main():
    for name, (rel, mesh, logs) in CASES.items():
        froude = os.path.join(REF, rel, "Froude")
        cg = os.path.join(froude, f"manning_{mesh}.cgns")
        x = [float(v) for v in h5(cg, "/Base/Zone/GridCoordinates/CoordinateX/ data")]
        y = [float(v) for v in h5(cg, "/Base/Zone/GridCoordinates/CoordinateY/ data")]
        z = [float(v) for v in h5(cg, "/Base/Zone/GridCoordinates/CoordinateZ/ data")]
        conn = [int(v) for v in h5(cg, "/Base/Zone/Elements/ElementConnectivity/ data", "%d")]
        mann = [float(v) for v in h5(cg, "/Base/Zone/Manning/{DataArray}/ data")]
        inflow = [int(v) for v in h5(cg, "/Base/Zone/ZoneBC/Inflow nodes/PointList/ data", "%d")]
        outflow = [int(v) for v in h5(cg, "/Base/Zone/ZoneBC/Outflow nodes/PointList/ data", "%d")]
        nnt, nelt = len(x), len(conn) // 3
        assert len(mann) == nelt
        with open(os.path.join(OUT, f"{name}.mesh"), "w") as fh:
            fh.write(f"{name}: nodes from manning_{mesh}.cgns\n{nnt}\n")
            for i in range(nnt):
                fh.write(f"{i + 1} {x[i]!r} {y[i]!r} {z[i]!r} 0.0\n")
            fh.write(f"elements\n{nelt}\n")
            for i in range(nelt):
                fh.write(f"{i + 1} {conn[3 * i]} {conn[3 * i + 1]} {conn[3 * i + 2]} {mann[i]!r}\n")
            fh.write(f"inflow nodes\n{len(inflow)}\n" + "".join(f"{v}\n" for v in inflow))
            fh.write(f"outflow nodes\n{len(outflow)}\n" + "".join(f"{v}\n" for v in outflow))
            fh.write("wall nodes\n0\n")
        shutil.copy(cg, os.path.join(OUT, f"{name}.cgns"))
        shutil.copy(os.path.join(froude, f"boundary_table_manning_{mesh}.cgns"),
                    os.path.join(OUT, f"{name}.boundary_table"))
        shutil.copy(os.path.join(froude, "donnees.f"), os.path.join(OUT, f"{name}.donnees.f"))
        sol = os.path.join(REF, rel, "Solution")
        shutil.copy(os.path.join(sol, "Manning.txt"), os.path.join(OUT, f"{name}.Manning.txt"))
        shutil.copy(os.path.join(sol, "Entree_gen_manning.txt"),
                    os.path.join(OUT, f"{name}.Entree_gen_manning.txt"))
        expected = {"nnt": nnt, "nelt": nelt, "ndi": len(inflow), "ndo": len(outflow),
                    "source": f"{rel}/Froude/" + ", ".join(logs), "logs": {}}
        for lg in logs:
            expected["logs"][lg] = parse_log(os.path.join(froude, lg))
        with open(os.path.join(OUT, f"{name}.expected.json"), "w") as fh:
            json.dump(expected, fh, indent=1)
        print(name, nnt, nelt, len(inflow), len(outflow))
    shutil.copy(os.path.join(REF, "src/cuteflow/donnees.f"), os.path.join(OUT, "default.donnees.f"))
    shutil.copy(os.path.join(REF, "Calibration_manning/Exemple2/base_files/mesh_6072.cgns"),
                os.path.join(OUT, "mesh_6072.cgns"))


if __name__ == "__main__":
    main()
