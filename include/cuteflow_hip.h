/*
 * cuteflow_hip.h - C ABI of the MI355X-native St-Venant finite-volume step.
 *
 * The reference (CuteFlow, CUDA Fortran) has no operator API: its hot path is
 * a set of argument-less kernels that talk through module globals and are
 * launched by the Fortran host loop.  This header is the drop-in boundary
 * that replaces that loop body for one subdomain.  A Fortran host binds it
 * with iso_c_binding, Python with ctypes.  All pointers are plain host
 * pointers unless the name says _device; no ownership is taken (copy
 * semantics).  Indices in tables are 1-based, as in the reference.
 *
 * Status convention: 0 = ok; negative = error (see cf_error_string).
 * Threading: one host thread per handle; all device work of a handle is
 * ordered on the handle's HIP stream.
 */
#ifndef CUTEFLOW_HIP_H
#define CUTEFLOW_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CF_MAX_IO 20      /* inlets/outlets (global_data.cuf:54, dimension(20)) */
#define CF_MAX_HYDRO 64   /* points per inflow hydrograph */

/* cf_params.options bits.  By default the library stores cells in its own
 * order (a Hilbert curve of the centroids, so that each 256-cell tile is a
 * compact patch), skips tiles that are dry with dry neighbours and hands
 * the tiles holding water to the first workgroups of each step; all of this
 * is exact: per-cell results are identical bits in any order, with or
 * without skipping. */
#define CF_OPT_KEEP_ORDER 1      /* keep the caller's cell order on the device */
#define CF_OPT_NO_TILE_SKIP 2    /* run every tile every step */
#define CF_OPT_TILE_128 4        /* 128-cell tiles (workgroups) instead of 256 */
#define CF_OPT_TILE_64 8         /* 64-cell tiles (one wavefront) */
#define CF_OPT_XCD_CONTIG 16     /* each XCD (L2) runs contiguous runs of tiles */
#define CF_OPT_FIXED_TILE_ORDER 32 /* workgroups take the tiles in a fixed order instead of tiles
                                      holding water first (re-ranked every 32 steps) */

enum {
    CF_OK = 0,
    CF_ERR_ARG = -1,        /* invalid argument or unsupported option */
    CF_ERR_HIP = -2,        /* HIP runtime error */
    CF_ERR_DT = -3,         /* dt sanity failure (cfl_condition.cuf:129-132) or NaN state */
    CF_ERR_MESH = -4,       /* inconsistent mesh */
    CF_ERR_NOMEM = -5,
    CF_ERR_IO = -6,         /* file missing, unreadable, or not in the expected layout */
    CF_ERR_SIZE = -7        /* input larger than a no-allocation entry point's fixed scratch */
};

/* Every constant the reference kernels read from `constant` memory
 * (global_data_device.cuf, set by data_transfer.cuf:12-96).  String options
 * are passed as the integer codes of data_transfer.cuf:43-66. */
typedef struct cf_params {
    double gp;               /* GP */
    double tolisec;          /* wet/dry threshold */
    double tol;              /* friction velocity floor and loop tolerance (default 1e-6f) */
    double cfl;
    double ts;               /* end time; use 1e300 for a fixed step count */
    double constant_dt;      /* used when is_dt_constant == 1 */
    double override_manning; /* used when is_override_manning == 1 */
    double tol_reg_perm;     /* steady-state tolerance: > 0 stops a run at steady state (see cf_set_discharge) */
    int iflux;               /* 1 = HLLC, 2 = Riadh's state HLLC, 3 = WAF (Loukili & Soulaimani), 4 = MUSCL-Hancock */
    int ilimiteur;           /* WAF limiter 0..4; MUSCL slope limiter 0 none, 1 superbee, 2 van Leer, 3 minmod */
    int friction;            /* 0/1 */
    int fricimplic;          /* 0 explicit, 1 (I-dt/2 J)^-1, 2 (I-dt B)^-1 */
    int inlet;               /* 1 'transm', 2 'inflow', 3 'phys_inflow' (the inlet's discharge shared
                                over its faces in proportion to the depth; not with cf_adj_*) */
    int outlet;              /* 1 'fixedheight', 2 'transm' */
    int is_dt_constant;
    int is_override_manning;
    int nombre_entrees;
    int options;             /* CF_OPT_* bits (layout choices; results do not depend on them) */
    double debitglob[CF_MAX_IO];
    double h_sortie[CF_MAX_IO];
    double long_entree[CF_MAX_IO];
    /* inflow hydrograph (no reference counterpart; its debit_var/loi_debitglob
       are parsed but never used, global_data.cuf:48-49): per inlet hydro_n
       points (t, Q), linear in between, constant outside; 0 = debitglob. */
    int hydro_n[CF_MAX_IO];
    double hydro_t[CF_MAX_IO * CF_MAX_HYDRO];
    double hydro_q[CF_MAX_IO * CF_MAX_HYDRO];
} cf_params;

/* Mesh as the reference holds it after lecture_maillage (maillage.cuf:4-166). */
typedef struct cf_mesh {
    int nelt;                /* cells (triangles) */
    int nnt;                 /* nodes */
    const int *conn;         /* 3*nelt, connectivite(i,1:3), 1-based, cell-major */
    const double *coords;    /* 3*nnt, coordonnees(n,1:3) = x,y,z, node-major */
    const int *boundary;     /* 6*nelt, boundary(nelt,3,2) in Fortran column-major
                                order; self-references are turned into walls
                                (set_undefined_boudaries_as_walls) */
    const double *manning;   /* nelt per-cell n (manning_nelt), or NULL */
} cf_mesh;

typedef struct cf_handle cf_handle;

/* Builds boundary(nelt,3,2) (Fortran order) from the connectivity and the
 * inlet/outlet/wall node lists: replaces tab_bound_tri + boundary_kind
 * (maillage.cuf:781-1110) with an O(N log N) directed-edge match.  Self
 * references (undefined faces) are kept, as in the reference's
 * boundary_table file (maillage.cuf:55-68).  num_in/num_out may be NULL
 * (all 1). */
int cf_build_boundary(int nelt, int nnt, const int *conn,
                      int ndi, const int *ndinput, const int *num_in,
                      int ndo, const int *ndoutput, const int *num_out,
                      int ndw, const int *ndwall, int *boundary);

/* cf_build_boundary computed on the current GPU: the same table bit for bit
 * (host arrays in and out).  The directed edges are matched by a radix sort
 * and one binary search per face, the boundary classes by one thread per
 * face; for large meshes it replaces the host sort (and the reference's
 * all-pairs GPU search, maillage.cuf:881-952). */
int cf_build_boundary_device(int nelt, int nnt, const int *conn,
                             int ndi, const int *ndinput, const int *num_in,
                             int ndo, const int *ndoutput, const int *num_out,
                             int ndw, const int *ndwall, int *boundary);

/* cf_build_boundary_device with no device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  conn, the node lists and boundary are DEVICE pointers (same
 * layouts as above, allocated by the caller); the work is ordered on
 * `stream` (a hipStream_t, NULL for the null stream) and the call returns
 * when the table is complete.  Its working memory is one fixed array per
 * GPU, held until the process ends: at most 131072 cells and 131072 nodes
 * (larger meshes return CF_ERR_SIZE).  Every call on a GPU uses that GPU's
 * one array, so calls that overlap on one GPU (other host threads, other
 * streams) race; the calls themselves call no hipMalloc or hipFree. */
int cf_build_boundary_device_noalloc(int nelt, int nnt, const int *conn,
                                     int ndi, const int *ndinput, const int *num_in,
                                     int ndo, const int *ndoutput, const int *num_out,
                                     int ndw, const int *ndwall, int *boundary, void *stream);

/* Inlet lengths long_entree (maillage.cuf:1423-1452) from a boundary table. */
int cf_inlet_lengths(int nelt, int nnt, const int *conn, const double *coords,
                     const int *boundary, int nentrees, double *long_entree);

/* Creates a solver for one subdomain: copies the mesh in, computes geometry
 * on the device (compute_surface_and_bathy, maillage.cuf:1284-1331, the
 * cotemin loop :132-159, construct_coo_table_elemwise :1254-1282) and the
 * neighbour stencil (find_upwind_cells :954-978). */
int cf_create(const cf_mesh *mesh, const cf_params *params, cf_handle **out);
void cf_destroy(cf_handle *h);
const char *cf_error_string(int status);

/* State vdlg(3,nelt) (h,hu,hv per cell); copy in / copy out.  Setting the
 * state resets tc to 0 and recomputes dt (finite_volumes_method.cuf:113-116). */
int cf_set_state(cf_handle *h, const double *U);
int cf_get_state(cf_handle *h, double *U);
/* Same on device memory (3*nelt doubles) on the handle's stream. */
int cf_set_state_device(cf_handle *h, const double *dU);
int cf_get_state_device(cf_handle *h, double *dU);

/* Current dt (from the current state) and time. */
int cf_initial_dt(cf_handle *h, double *dt);
int cf_get_time(cf_handle *h, double *tc, double *dt);

/* Runs up to nsteps steps of the loop body of finite_volumes_method.cuf:123-242
 * (flux, source, update, cfl/dt, state swap, tc += dt); stops early when
 * tc >= ts - tol or on a dt/NaN failure.  dt_hist (nsteps+1, may be NULL)
 * receives the dt used by each step and then the next dt; *steps_done the
 * number of steps run.  Returns CF_ERR_DT on a dt failure. */
int cf_step(cf_handle *h, int nsteps, double *dt_hist, int *steps_done);
/* cf_step split in two: cf_step_async only enqueues the steps on the
 * handle's stream (no host synchronisation); cf_step_finish waits for them
 * and returns what cf_step returns.  Every cf_step_async must be followed by
 * one cf_step_finish before any other call on the handle. */
int cf_step_async(cf_handle *h, int nsteps);
int cf_step_finish(cf_handle *h, double *dt_hist, int *steps_done);
/* Readies the handle for runs of nsteps steps without running any: sizes the
 * per-run history and captures the run's launches in a graph, the one-time
 * work a first cf_step/cf_step_async of that length would do (no reference
 * counterpart: the reference launches each step's kernels from the host).
 * Call it outside a timed region or before a latency-bound loop. */
int cf_step_prepare(cf_handle *h, int nsteps);

/* Work done by the last cf_step/cf_step_finish: tiles that ran (the others
 * were dry with dry neighbours and skipped), tiles x steps in total, and the
 * cells per tile. */
int cf_last_run_stats(cf_handle *h, long long *tiles_run, long long *tiles_total, int *tile_cells);

/* Inflow discharges debitglob(1:n) and hydrographs. */
int cf_set_inflow(cf_handle *h, const double *debitglob, int n);
int cf_set_hydrograph(cf_handle *h, int inlet, int npts, const double *t, const double *q);
/* Per-cell Manning n (manning_nelt, nelt doubles in the caller's order; host
 * or device pointer). */
int cf_set_manning(cf_handle *h, const double *n);

/* Geometry computed on the device: surf, zm, cotemin_arr (each nelt, may be
 * NULL) and cotemin = min(cotemin_arr). */
int cf_get_geometry(cf_handle *h, double *surf, double *zm, double *cotemin_arr,
                    double *cotemin);

/* Stream of the handle (hipStream_t as void*) and its replacement, so that a
 * caller (e.g. PyTorch) can order the solver on its own stream. */
void *cf_get_stream(cf_handle *h);
int cf_set_stream(cf_handle *h, void *stream);

/* ---- Batches: many independent subdomains in one handle ----------------
 *
 * No reference counterpart: the reference runs one mesh per process
 * (finite_volumes_method.cuf:15-257).  A batch handle holds ninst instances
 * (meshes of any sizes, e.g. ensemble members or calibration candidates),
 * each with its own parameters, state, dt, time and stop/error status; one
 * launch per step advances all of them, and each instance's results are
 * bit-identical to a single-instance handle on the same inputs.  All
 * instances must share iflux, friction, fricimplic and options (one kernel
 * variant serves the batch); other fields may differ.  The single-instance
 * functions above act on instance 0.  Instance indices are 0-based. */
int cf_batch_create(int ninst, const cf_mesh *meshes, const cf_params *params, cf_handle **out);
int cf_batch_size(cf_handle *h);
int cf_batch_set_state(cf_handle *h, int inst, const double *U);
int cf_batch_get_state(cf_handle *h, int inst, double *U);
int cf_batch_set_state_device(cf_handle *h, int inst, const double *dU);
int cf_batch_get_state_device(cf_handle *h, int inst, double *dU);
int cf_batch_get_time(cf_handle *h, int inst, double *tc, double *dt);
/* Runs up to nsteps steps of every instance that has not stopped; an
 * instance stops at its own ts or dt failure while the others go on.
 * dt_hist (ninst x (nsteps+1), row per instance, may be NULL) as for
 * cf_step; steps_done and status (ninst each, may be NULL) per instance,
 * status CF_OK or CF_ERR_DT.  Returns CF_ERR_DT if any instance failed.
 * cf_step_async + cf_batch_step_finish is the split form. */
int cf_batch_step(cf_handle *h, int nsteps, double *dt_hist, int *steps_done, int *status);
int cf_batch_step_finish(cf_handle *h, double *dt_hist, int *steps_done, int *status);
int cf_batch_set_inflow(cf_handle *h, int inst, const double *debitglob, int n);
int cf_batch_set_hydrograph(cf_handle *h, int inst, int inlet, int npts, const double *t, const double *q);
int cf_batch_set_manning(cf_handle *h, int inst, const double *n);
int cf_batch_get_geometry(cf_handle *h, int inst, double *surf, double *zm, double *cotemin_arr,
                          double *cotemin);

/* Discharge bookkeeping and the steady-state stop (flux_riemann.cuf:125-126,
 * 161 record the inflow and outflow edge discharges; check_steady_state,
 * finite_volumes_method.cuf:370-393).  With it on, every step records its
 * flux's inflow discharge (inflow edges, inlets 'inflow' and 'phys_inflow') and outflow
 * discharge (fixed-level outlet edges), each summed over the cells in the
 * caller's order, and a run stops after a step that makes more than 1000
 * consecutive steps with Qin > 1e-8f and |Qin - Qout| / Qin < tol_reg_perm
 * (status CF_OK, steps_done < nsteps).  It is on whenever an instance has
 * tol_reg_perm > 0; cf_set_discharge(h, 1) turns it on for the discharges
 * alone.  cf_batch_get_discharge copies the last run's per-step Qin and
 * Qout (steps_done entries each, may be NULL) and whether the instance
 * stopped at steady state. */
int cf_set_discharge(cf_handle *h, int on);
int cf_batch_get_discharge(cf_handle *h, int inst, double *q_in, double *q_out, int *steady);
int cf_get_discharge(cf_handle *h, double *q_in, double *q_out, int *steady);

/* Schedule autotuner.  No reference counterpart (the reference's launch
 * shapes are fixed: 24-thread flux blocks, flux_riemann.cuf; 1024-thread
 * reductions, cfl_condition.cuf:75).  Times nsteps steps of the batch
 * (meshes, params as for cf_batch_create; states: ninst host pointers to
 * (3, nelt) initial states) under each of CF_TUNE_NCAND schedules and
 * returns in *best_options params[0].options with the tuned bits
 * (CF_OPT_TILE_128, CF_OPT_TILE_64, CF_OPT_NO_TILE_SKIP, CF_OPT_XCD_CONTIG)
 * set to the fastest (a candidate that skips dry tiles is taken over a
 * faster one that does not if it is within 3 %: the skip also shortens the
 * replay and reverse sweep of derivative runs, which are not timed); pass it
 * as options to cf_batch_create.  Candidate k
 * has cells per tile 256 >> (k / 4) (256, 128, 64), dry-tile skipping off
 * if k & 2, XCD-contiguous tiles if k & 1.  times_ms (CF_TUNE_NCAND
 * entries, may be NULL) receives each candidate's best time of the timed
 * runs, or -1 if it was not run (a failure of that candidate).  The results
 * of a run do not depend on the schedule: only its speed does. */
#define CF_TUNE_NCAND 12
int cf_batch_autotune(int ninst, const cf_mesh *meshes, const cf_params *params, const double *const *states,
                      int nsteps, int *best_options, double *times_ms);

/* ---- Discrete adjoint and tangent-linear model --------------------------
 *
 * No reference counterpart: the reference calibrates Manning n with
 * repeated forward runs (Calibration_manning/).  These entry points give
 * exact derivatives of the discrete time step (wet/dry switches, wave-speed
 * cases, limiter regions and dt frozen as taken in the forward) of the
 * water levels h + zm at gauge cells and of the per-cell maximum depth, with
 * respect to per-cell Manning n and the initial state.  Per-cell arrays
 * (nelt or 3*nelt doubles, caller's cell order) may be host or device
 * pointers; sample arrays (nsamp x ngauges, row per sample) are host
 * pointers.  Requires is_override_manning == 0 in every instance. */

/* Derivative options (bits; 0 = the exact derivative of the discrete step).
 * CF_ADJ_FREEZE_LIMITERS: the WAF limiter values are taken as constants, like
 * the wet/dry switches and dt.  Where two face depths agree to rounding (water
 * at rest, uniform flow) the limiter ratio du / (dl + 1e-16)
 * (flux_riemann.cuf:461-463) is rounding noise with derivatives of order
 * 1e16, and over thousands of steps the exact derivative of the WAF step
 * grows without bound; frozen limiters give the derivative of the step with
 * the limiter values of the forward.  The forward is unchanged. */
#define CF_ADJ_FREEZE_LIMITERS 1
int cf_adj_set_flags(cf_handle *h, int flags);

/* Gauge cells of instance inst (0-based, caller's order). */
int cf_adj_set_gauges(cf_handle *h, int inst, int n, const int *cells);
/* Recorded forward: advances every instance by exactly nsteps steps (an
 * instance that stops earlier is an error) from its current state, keeping
 * dt and tc of every step, a checkpoint every checkpoint_every steps, the
 * gauge levels after every sample_every-th step (nsamp = nsteps /
 * sample_every samples) and, when want_hmax, the per-cell maximum depth over
 * those samples. */
int cf_adj_forward(cf_handle *h, int nsteps, int sample_every, int checkpoint_every, int want_hmax);
int cf_adj_num_samples(cf_handle *h);
int cf_adj_get_samples(cf_handle *h, int inst, double *eta, double *hmax);
/* Cotangents of the outputs of instance inst: gauge levels (nsamp x ng),
 * hmax (nelt) and the final state (3*nelt); NULL = 0. */
int cf_adj_set_seed(cf_handle *h, int inst, const double *g_eta, const double *g_hmax, const double *g_state);
/* Reverse sweep over the recorded run: segments of checkpoint_every steps
 * replayed from their checkpoints, then reversed step by step (the first
 * sweep after a record takes the last segment as the recorded forward wrote
 * it; a later sweep over the same record replays it too).  It takes
 * the Manning n of the recorded run, whatever n was set since (the
 * tangent-linear run takes the current n); after an inflow or hydrograph
 * change since the record it returns CF_ERR_ARG (record again). */
int cf_adj_backward(cf_handle *h);
/* Gradient of instance inst with respect to per-cell Manning n (nelt) and
 * to the initial state (3*nelt); either may be NULL. */
int cf_adj_get_gradient(cf_handle *h, int inst, double *g_manning, double *g_state0);
/* Tangent-linear run over the recorded run along (dn, dU0) (NULL = 0),
 * from the recorded initial state with the recorded dt and tc at the current
 * Manning n: the gauge levels (eta) and final state (state) of that run (at
 * the recorded n: the recorded forward's bits) and their directional
 * derivatives, and that of hmax. */
int cf_tlm_set_direction(cf_handle *h, int inst, const double *dn, const double *dU0);
int cf_tlm_run(cf_handle *h);
int cf_tlm_get(cf_handle *h, int inst, double *eta, double *d_eta, double *d_hmax, double *d_state,
               double *state);
/* Branch decisions of tangent-linear runs, for finite-difference checks
 * (no reference counterpart).  mode 1: the next cf_tlm_run records, per
 * step and cell, whether the update left the cell wet (the clamp of
 * finite_volumes_method.cuf:297-303 not taken); mode 2: the next runs count
 * per cell the steps whose decision differs from the recorded run (a run at
 * perturbed n that wets or dries a cell at another step).  Modes 3 and 4 do
 * the same and also record / compare a signature of every branch of the
 * cell's update (dry-face thresholds, wave-speed cases, solver regions,
 * max/min selections, the clamp): 4 bytes per step and cell.  0: off.
 * Needs a recorded forward; modes 2 and 4 need masks recorded for its step
 * count; a new cf_adj_forward turns them off.  cf_tlm_get_flips returns the
 * counts of instance inst (kind 0: wet/dry clamp, 1: any branch) in the
 * caller's cell order and zeroes them if reset. */
int cf_tlm_set_mask_mode(cf_handle *h, int mode);
int cf_tlm_get_flips(cf_handle *h, int inst, int kind, int *flips, int reset);

/* Elementwise pow used by the friction kernel, evaluated on the device for
 * n values (x > 0): lets tests compare the device pow bit-for-bit. */
int cf_debug_pow(int n, const double *x, double y, double *out);

/* The internal layout cf_create builds for a mesh (cell order, neighbour
 * codes, tile adjacency, inflow tiles, inflow/outflow cells), from the
 * GPU preparation (device = 1, what cf_create uses) or the host one
 * (device = 0): lets tests compare the two.  Arrays: perm (nelt), nbr
 * (4*nelt), conn (3*nelt), tile_off (ntiles+1), tile_adj (adj_cap, count in
 * *n_adj), tile_always (ntiles), io_cells (nelt, count in *n_io). */
int cf_debug_prepare(const cf_mesh *mesh, const cf_params *params, int tile, int keep_order, int device,
                     int *perm, int *nbr, int *conn, int *tile_off, int *tile_adj, int adj_cap, int *n_adj,
                     unsigned char *tile_always, int *io_cells, int *n_io);

#ifdef __cplusplus
}
#endif
#endif
