/* CGNS mesh and solution files, read and written in the CGNS library's HDF5
 * layout with the HDF5 C API (the CGNS library itself is not needed).  Host
 * code only, in its own shared library (libcuteflow_cgns.so), so that the
 * solver library carries no HDF5 dependency.
 *
 * Every function returns a cf status (CF_OK, or CF_ERR_ARG / CF_ERR_MESH /
 * CF_ERR_IO / CF_ERR_NOMEM); cf_cgns_error() gives the message of the last
 * failure on the calling thread.  Node ids are 1-based as in the files. */
#pragma once

#include "cuteflow_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* What cf_cgns_read_mesh will return for a file. */
typedef struct cf_cgns_info {
    int nnt;          /* vertices of the zone read (zone sizes(1)) */
    int nelt;         /* cells of the zone read (zone sizes(2)) */
    int ndi;          /* inflow node-list length: points of the BCInflow BCs with more than one point */
    int ndo;          /* outflow node-list length: likewise for BCOutflow */
    int nzones;       /* Zone_t nodes of the first base */
    int zone;         /* zone read, 1-based: the first whose Ordinal is 0 (no Ordinal node reads 0) */
    int has_manning;  /* the zone holds UserDefinedData "Manning" */
    int partitioned;  /* the zone holds ElemsToSend / ElemsToRecv (one part of a multi-rank mesh) */
} cf_cgns_info;

/* Sizes of the mesh in `path` (lec_mail_cgns, maillage.cuf:440-556: first
 * base, zone by ordinal, its sizes, and the first pass over the BCs). */
int cf_cgns_mesh_info(const char *path, cf_cgns_info *info);

/* Reads the mesh as lec_mail_cgns (maillage.cuf:398-778) does on one rank:
 *   coords   (nnt x 3, row-major x y z) from GridCoordinates/CoordinateX|Y|Z;
 *   conn     (nelt x 3, row-major) from the zone's first Elements_t section,
 *            which must be one TRI_3 section over the zone's nelt cells;
 *   ndinput / ndoutput (info.ndi / info.ndo entries): the point lists of the
 *            BCInflow / BCOutflow BCs with more than one point, in the
 *            file's BC order;
 *   numero_ndinput / numero_ndoutput: for the points of each BC, the 1-based
 *            position of the BC's name among inlet_name(1..nombre_entrees) /
 *            outlet_name(1..nombre_sorties) (the last match), else 1 with
 *            the reference's "not found" message;
 *   manning  (nelt, may be NULL): the first DataArray of UserDefinedData
 *            "Manning" unless is_override_manning == 1; zeros when not read.
 * inlet_name / outlet_name: nombre_entrees / nombre_sorties fixed-width
 * fields of name_len bytes (Fortran CHARACTER(name_len) arrays: blank or
 * NUL padded).  verbose prints the reference's console messages of this
 * routine.  A partitioned zone (ElemsToSend/ElemsToRecv) is refused with
 * CF_ERR_MESH: its ghost cells need the multi-rank halo exchange. */
int cf_cgns_read_mesh(const char *path, int nombre_entrees, const char *inlet_name, int nombre_sorties,
                      const char *outlet_name, int name_len, int is_override_manning, int verbose,
                      double *coords, int *conn, double *manning, int *ndinput, int *numero_ndinput,
                      int *ndoutput, int *numero_ndoutput);

/* Writes a new mesh file in the layout the reader takes and the reference's
 * converter produces (cute_to_cgns.cpp:118-160): base "Base" (cell dim 2,
 * physical dim 3), unstructured zone "Zone", TRI_3 section "Elements",
 * BCInflow "Inflow nodes" (one BC per inlet number, "Inflow nodes <k-1>",
 * when nombre_entrees > 1), BCOutflow "Outflow nodes" (likewise per outlet
 * number when nombre_sorties > 1), BCWall "Wall nodes", and, when manning
 * is not NULL, UserDefinedData "Manning" at CellCenter holding
 * "{DataArray}" (gen_manning.py:71-74).  numero_* may be NULL (all 1). */
int cf_cgns_write_mesh(const char *path, int nnt, int nelt, const double *coords, const int *conn,
                       int nombre_entrees, int ndi, const int *ndinput, const int *numero_ndinput,
                       int nombre_sorties, int ndo, const int *ndoutput, const int *numero_ndoutput,
                       int ndw, const int *ndwall, const double *manning);

/* cgns_export_init (pre_post_traitement.cuf:223-285) on an existing mesh
 * file (the reference's out_<meshfile> copy): BaseIterativeData
 * "TimeIterValues" with nsnap steps and TimeValues = times[0..nsnap-1], and
 * the zone's ZoneIterativeData "FlowSolutionPointers" naming Solution_0 ..
 * Solution_<nsnap-1> (32 characters each, blank padded).  Nodes of the same
 * name are replaced. */
int cf_cgns_export_init(const char *path, int nsnap, const double *times);

/* cgns_export (pre_post_traitement.cuf:287-399): writes FlowSolution
 * "Solution_<k>" at CellCenter on the zone, with the fields, in this order,
 * Water Depth (h), Bathymetry (zm), Free surface elevation
 * (h + zm - sol_z_offset), Manning (only when manning is not NULL; the
 * reference writes it unless is_override_manning), MomentumX/Y (hu, hv where
 * h > tolisec, else 0), MomentumZ (0), VelocityX/Y (hu/h, hv/h where
 * h > tolisec, else 0), VelocityZ (0).  U is nelt x 3 (h, hu, hv), the
 * caller's cell order. */
int cf_cgns_export(const char *path, int k, int nelt, const double *U, const double *zm, const double *manning,
                   double tolisec, double sol_z_offset);

/* Message of the last failing call on this thread ("" if none). */
const char *cf_cgns_error(void);

#ifdef __cplusplus
}
#endif
