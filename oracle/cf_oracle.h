/*
 * CPU oracle for the St-Venant finite-volume time step (TEST INFRASTRUCTURE).
 *
 * This is a plain-C restatement of the reference CUDA Fortran kernels of
 * CuteFlow, used only as the parity checker by tests/, by the smoke check and
 * by bench.py's CPU-baseline leg.  The product (cuteflow_amd/, the HIP library)
 * never links, imports or calls anything under oracle/.
 *
 * Conventions follow the reference's Fortran view:
 *   - cell and node ids stored in tables are 1-based;
 *   - the state U is (3, nelt) column-major, i.e. h,hu,hv per cell (AoS);
 *   - boundary table column 1 (b1) holds the neighbour id (>0) or the
 *     boundary class (-1 inflow, -2 outflow, -3 wall); column 2 (b2) holds
 *     the inlet/outlet number, b1/b2 are stored cell-major: b1[3*i + face].
 *   - upw[6*i + 2*face + s] is upwind_cells(s+1, face+1, i+1).
 */
#ifndef CF_ORACLE_H
#define CF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_IO 20
#define OR_MAX_HYDRO 64

typedef struct or_params {
    double gp;            /* gravity */
    double tolisec;       /* wet/dry threshold */
    double tol;           /* velocity floor in friction, loop tolerance */
    double cfl;
    double ts;            /* end time (use a huge value for a fixed step count) */
    double constant_dt;
    double override_manning;
    double tol_reg_perm;  /* steady-state relative tolerance (<=0 disables) */
    int iflux;            /* 1 HLLC, 3 WAF */
    int ilimiteur;        /* 0..4 */
    int friction;         /* 0/1 */
    int fricimplic;       /* 0,1,2 */
    int inlet;            /* 1 transm, 2 inflow, 3 phys_inflow */
    int outlet;           /* 1 fixedheight, 2 transm */
    int is_dt_constant;
    int is_override_manning;
    int nombre_entrees;
    int pad0;
    double debitglob[OR_MAX_IO];
    double h_sortie[OR_MAX_IO];
    double long_entree[OR_MAX_IO];
    /* inflow hydrograph extension: per inlet, hydro_n[k] points (t, Q);
       hydro_n[k] == 0 means the constant debitglob[k]. */
    int hydro_n[OR_MAX_IO];
    double hydro_t[OR_MAX_IO * OR_MAX_HYDRO];
    double hydro_q[OR_MAX_IO * OR_MAX_HYDRO];
    /* phys_inflow: wetted area sum(length x depth) of each inlet for the
       current state; or_run recomputes it before every step */
    double surf_entree[OR_MAX_IO];
} or_params;

/* project pow for x > 0 (exp(y*log(x)) in double-double), shared algorithm
 * with the device code */
double or_pow(double x, double y);

/* ---- mesh preprocessing (maillage.cuf) ---- */
void or_neighbours_allpairs(int nelt, const int *conn, int *b1, int *b2);
void or_neighbours_sorted(int nelt, const int *conn, int *b1, int *b2);
void or_boundary_kind(int nelt, const int *conn, int *b1, int *b2,
                      int ndi, const int *ndinput, const int *num_in,
                      int ndo, const int *ndoutput, const int *num_out,
                      int ndw, const int *ndwall);
void or_set_walls(int nelt, int *b1);
void or_upwind_cells(int nelt, const int *b1, int *upw);
void or_geometry(int nelt, const int *conn, const double *x, const double *y,
                 const double *z, double *vx, double *vy, double *xc, double *yc,
                 double *surf, double *zm, double *cotemin_arr, double *cotemin);
void or_long_entree(int nelt, const int *conn, const double *x, const double *y,
                    const int *b1, const int *b2, int nentrees, double *long_entree);

/* ---- initial conditions (initial_condition.cuf) ---- */
void or_ic_dam(int nelt, const double *xc, const double *yc, const double *zm,
               double tolisec, double x1, double y1, double x2, double y2,
               double h_amont, double u_amont, double v_amont,
               double h_aval, double u_aval, double v_aval, double *U);

/* ---- per-step kernels ---- */
double or_inflow_q(const or_params *p, int id, double tc);
void or_compute_flux(const or_params *p, int nelt, const int *conn,
                     const double *vx, const double *vy, const double *xc,
                     const double *yc, const double *zm, const int *b1,
                     const int *b2, const int *upw, const double *U0,
                     double dt, double tc, double *resc, double *qout_cell, double *qin_cell);
void or_source_terms(const or_params *p, int nelt, const double *vx,
                     const double *vy, const double *surf, const double *zm,
                     const double *manning, const int *b1, const double *U0,
                     double dt, double *sb, double *sf, double *afm1);
void or_update(const or_params *p, int nelt, const double *surf,
               const double *U0, const double *resc, const double *sb,
               const double *sf, const double *afm1, double dt, double *U);
double or_cfl_dt(const or_params *p, int nelt, const double *cotemin_arr,
                 const double *U);

/* ---- the time loop (finite_volumes_method.cuf:113-242) ----
 * Runs at most nsteps steps (or until tc >= ts - tol, an error, or steady
 * state).  U (3*nelt) is updated in place.  dt_hist[k] receives the dt used
 * by step k (k < steps done), dt_hist[steps_done] the dt for the next step.
 * qout_hist / qin_hist (optional) receive the outflow / inflow discharge of
 * each step's flux, summed over cells in cell order.  *steady (optional) is
 * set to 1 when the run stopped at steady state (check_steady_state,
 * finite_volumes_method.cuf:370-393).
 * Returns 0 on success, -1 on a dt sanity failure (dt>9999, dt<1e-12 or NaN).
 */
/* calcul_surf_entree: wetted area of each inlet (length x depth summed over
 * its cells in cell order, a cell's last face on the inlet), for phys_inflow */
void or_surf_entree(const or_params *p, int nelt, const double *vx, const double *vy, const int *b1,
                    const int *b2, const double *U, double *surf_entree);

int or_run(const or_params *p, int nelt, const int *conn, const double *vx,
           const double *vy, const double *xc, const double *yc,
           const double *surf, const double *zm, const double *cotemin_arr,
           const double *manning, const int *b1, const int *b2, const int *upw,
           double *U, int nsteps, double *dt_hist, double *tc_out,
           int *steps_done, double *qout_hist, double *qin_hist, int *steady, int nthreads);

/* face solvers (flux in the global frame) and limiter, for the pin tests */
void or_riemann_face(int iflux, int ilimiteur, double gp, double n1, double n2,
                     double hl, double ul, double vl, double hr, double ur, double vr,
                     double hl_upw, double vl_upw, double hr_upw, double vr_upw,
                     double d_mrl, double d_mkvol, double d_mkvoisin, double dt,
                     double flux[3]);
double or_limiter_value(int ilimiteur, double r, double c);
double or_slope_limiter_value(int ilimiteur, double r1, double r2);
/* solver_Riemann_muscl (iflux 4) with its upwind values, global frame */
void or_muscl_face(int ilimiteur, double gp, double tolisec, double n1, double n2,
                   double hl, double ul, double vl, double hr, double ur, double vr,
                   double hl_upw, double ul_upw, double vl_upw,
                   double hr_upw, double ur_upw, double vr_upw,
                   double d_mrl, double d_mkvol, double d_mkvoisin, double dt, double flux[3]);

#ifdef __cplusplus
}
#endif
#endif
