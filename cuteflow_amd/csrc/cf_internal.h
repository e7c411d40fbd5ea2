// Shared definitions of the HIP translation units: device data layout,
// per-instance parameters and control block, and the solver handle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/cuteflow_hip.h"

namespace cf {

// cells per tile = threads per workgroup of the step kernel
constexpr int kTile = 256;
constexpr int kWaves = kTile / 64;

// geometry arrays (structure of arrays, stride = padded cell count)
enum { G_N1 = 0, G_N2 = 3, G_LN = 6, G_SURF = 9, G_COT = 10, G_XC = 11, G_YC = 12, G_COUNT = 13 };

// Per-instance constants (the reference's `constant` memory,
// global_data_device.cuf, set by data_transfer.cuf:12-96).
struct KParams {
    double gp, tolisec, tol, cfl, ts, constant_dt, override_manning;
    double pad_dt;  // min(1e6, 1e5 if nelt % 1024 != 0): the reduction's padding
    int ilimiteur, inlet, outlet, is_dt_constant, is_override_manning;
    int hydro_off;  // offset of this instance's hydrograph tables in the hydro array
    double debitglob[CF_MAX_IO];
    double h_sortie[CF_MAX_IO];
    double long_entree[CF_MAX_IO];
    int hydro_n[CF_MAX_IO];
};

// Per-instance control block, resident on the device across runs.
struct InstCtrl {
    double dt;        // dt of the current state (used by the next step)
    double tc;        // time of the current state
    int status;       // 0 running, 1 ts reached, -1 dt/NaN failure
    int steps;        // steps done in the current run
    unsigned gstep;   // steps done since creation (dry-tile stamps)
    int par;          // state buffer holding the current state at the run's start
};

struct StepArgs {
    double4 *st[2];                      // ping-pong state buffers (h, hu, hv, zm)
    const int4 *__restrict__ nbr;        // neighbour codes, global cell indices
    const double *__restrict__ geo;      // G_COUNT x ncell
    const double4 *__restrict__ dfc;     // centroid distances across faces
    const double *__restrict__ manning;  // per cell
    const double *__restrict__ hydro;    // per instance: t[20*64] then q[20*64]
    const KParams *__restrict__ kp;      // per instance
    InstCtrl *ctrl;                      // per instance
    double *dt_hist;                     // [inst*hist_stride + k]
    double *tc_hist;
    int *ran;                            // tiles that ran, [inst*hist_stride + k]
    double *partial;                     // per tile
    unsigned *stamp;                     // per tile
    const int *__restrict__ tile_adj_off;
    const int *__restrict__ tile_adj;
    const unsigned char *__restrict__ always;
    const int *__restrict__ tile_inst;   // instance of each tile
    const int *__restrict__ inst_end;    // per instance: one past its last real cell
    const int *__restrict__ inst_tile;   // per instance: first tile; [ninst] = total
    // explicit buffers (adjoint recompute): when set, the step reads in and writes out
    const double4 *in_x;
    double4 *out_x;
    int ncell;        // padded cell count (geo stride)
    int hist_stride;
    int step;         // index of this step in the run
    int skip;         // dry-tile skipping enabled
};

}  // namespace cf

struct cf_handle {
    int ninst = 0;
    int ncell = 0;   // padded total
    int ntiles = 0;
    std::vector<int> nelt, nnt, cell_off, tile_off;    // per instance (tile_off has ninst+1)
    std::vector<cf_params> prm;
    std::vector<cf::KParams> kp;
    std::vector<std::vector<int>> hperm;               // per instance: internal -> caller (local)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int iflux = 1, fric = 0, skip = 1;
    // device arrays
    double4 *st[2] = {nullptr, nullptr};
    int4 *nbr = nullptr;
    double *geo = nullptr;
    double4 *dfc = nullptr;
    double *manning = nullptr;
    double *hydro = nullptr;
    double *zm = nullptr;
    double *stage = nullptr;           // 3 * max cells of an instance
    int *perm = nullptr;               // per cell: caller's local index
    cf::KParams *d_kp = nullptr;
    cf::InstCtrl *ctrl = nullptr;
    double *partial = nullptr;
    unsigned *stamp = nullptr;
    int *tile_adj_off = nullptr, *tile_adj = nullptr;
    unsigned char *always = nullptr;
    int *tile_inst = nullptr, *inst_end = nullptr, *inst_tile = nullptr;
    double *dt_hist = nullptr, *tc_hist = nullptr;
    int *ran = nullptr;
    int cap = 0;                       // steps per run the history arrays hold
    // host mirror of the control block after the last finished run
    std::vector<cf::InstCtrl> hctrl;
    std::vector<long long> last_run, last_total;
    int pending = -1;
    int max_cells = 0;
    // graphs of whole runs, keyed by (steps, parity pattern)
    struct Graph {
        int nsteps;
        std::vector<int> par;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    int use_graphs = 1;
};
