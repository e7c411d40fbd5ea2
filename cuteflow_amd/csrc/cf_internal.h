// Shared definitions of the HIP translation units: device data layout,
// per-instance control block, kernel arguments, the solver handle and small
// host helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/cuteflow_hip.h"
#include "cf_cell.h"

namespace cf {

// cells per tile = threads per workgroup of the per-cell kernels: 256 by
// default, 128 or 64 by option (h->tile); kernels are compiled for the
// largest and read the tile size from blockDim.x
constexpr int kTile = 256;
constexpr int kWaves = kTile / 64;

// dt slots per instance (spread the atomics of a step over 8 addresses),
// each on its own 128-byte line
constexpr int kDtSlots = 8;
constexpr int kSlotStride64 = 16;
constexpr int kSlotStride32 = 32;
// bit pattern of +inf: no tile has folded its minimum in yet
constexpr unsigned long long kDtSlotEmpty = 0x7ff0000000000000ull;

// Tile order of the step kernel: every kReorderEvery steps of a run (and at
// its start) the tiles are ranked by what they did at the last step that
// ran them: held water, ran with only dry cells, skipped.
constexpr int kReorderEvery = 32;
enum { kTileWet = 0, kTileDry = 1, kTileSkipped = 2 };
// replay actions per tile and step
enum { kReplayKeep = 0, kReplayRun = 1, kReplayRewrite = 2 };

// Per-instance control block, resident on the device across runs.
struct InstCtrl {
    double dt;        // dt of the current state (used by the next step)
    double tc;        // time of the current state
    int status;       // 0 running, 1 ts reached, 2 steady state, -1 dt/NaN failure
    int steps;        // steps done in the current run
    unsigned gstep;   // steps done since creation (dry-tile stamps)
    int par;          // state buffer holding the current state at the run's start
    int reg;          // consecutive steps with |Qin - Qout| / Qin < tol_reg_perm
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
    // dt reduction of a step: every tile that runs folds its minimum into
    // one of kDtSlots slots of its instance (atomic min on the bits, exact)
    // and counts itself; a tile with a NaN state raises the instance's flag
    unsigned long long *dtslot;          // [inst * kDtSlots * kSlotStride64 + slot * kSlotStride64]
    unsigned *cntslot;                   // [inst * kDtSlots * kSlotStride32 + slot * kSlotStride32]
    unsigned *nanflag;                   // [inst * kSlotStride32]
    unsigned *stamp;                     // per tile
    const int *__restrict__ tile_adj_off;
    const int *__restrict__ tile_adj;
    const unsigned *__restrict__ always;   // per tile: 1 = always runs (inflow faces); 32-bit for a scalar load
    const int *__restrict__ tile_inst;   // instance of each tile
    const int *__restrict__ inst_end;    // per instance: one past its last real cell
    const int *__restrict__ inst_tile;   // per instance: first tile; [ninst] = total
    // Replay of a recorded run (reverse sweep): when dt_fix is set the step
    // takes dt and tc of step `step` from dt_fix/tc_fix, reads in_x, writes
    // out_x and the residual x to xs (3 x ncell), and skips the dt reduction.
    // A recorded forward that also keeps the tape of its steps (out_x set,
    // dt_fix null) writes the states after the step to out_x and the
    // residuals to xs, as the replay would.
    const double *dt_fix;
    const double *tc_fix;
    const double4 *in_x;
    double4 *out_x;
    double *xs;
    const double *zm;  // bed elevation per cell (the w of every state)
    // Tile activity of a recorded run, row per step, one byte per tile: the
    // recorded forward writes whether each tile ran.  In a replay a tile
    // that did not run holds dry cells with dry neighbours, whose state
    // after the step is (+0, +0, +0, zm), as before it.  out_x is a reused
    // buffer, so the replay reads one action byte per tile (replay_act,
    // kReplay*): run the tile, rewrite its slot to (+0, +0, +0, zm) (the slot
    // last held a step at which the tile ran, or an unknown one), or nothing.
    unsigned char *tile_ran;
    const unsigned char *replay_act;
    // phys_inflow: per instance, CF_MAX_IO wetted inlet areas of the step's
    // input state (null when no instance uses phys_inflow)
    const double *surf_in;
    int ncell;        // padded cell count (geo stride)
    int tile_cells;   // cells per tile = workgroup size
    int hist_stride;
    int step;         // index of this step in the run (absolute in a replay)
    int skip;         // dry-tile skipping enabled
    int xcd;          // workgroup -> tile order: 0 identity, 1 XCD-contiguous
    // Workgroup -> (tile, instance) of the step kernel.  With the wet-first
    // order the step kernel also records each tile's class (kTileWet,
    // kTileDry, kTileSkipped) in tile_class, and reorder_kernel rebuilds the
    // map from it: wet tiles take the first workgroups, then the dry tiles
    // that run, then the skipped ones.  Any permutation gives the same bits.
    const int2 *__restrict__ tile_map;
    unsigned char *tile_class;  // null: fixed order, classes not recorded
};

// Tile processed by this workgroup.  Workgroups are dispatched round-robin
// over the 8 XCDs (each with its own L2); with xcd set, XCD x gets one
// contiguous run of tiles instead of every 8th tile, so the neighbour
// reads across tile edges stay in that XCD's L2.  A bijection on
// [0, gridDim.x) for any grid size.
__device__ inline int tile_of_block(int xcd) {
    const int bid = blockIdx.x;
    if (!xcd) return bid;
    const int g = gridDim.x, x = bid & 7, q = g >> 3, r = g & 7;
    return x * q + (x < r ? x : r) + (bid >> 3);
}

// Device loader of the cell body: states, neighbours, geometry.
struct DevLd {
    const double4 *st_;
    const int4 *nbr_;
    const double4 *dfc_;
    const double *geo_;
    const double *man_;
    int stride;
    __device__ __forceinline__ St<double> state(int c) const {
        const double4 v = st_[c];
        return {v.x, v.y, v.z, v.w};
    }
    __device__ __forceinline__ I4 nbr(int c) const {
        const int4 v = nbr_[c];
        return {v.x, v.y, v.z, v.w};
    }
    __device__ __forceinline__ D4 dfc(int c) const {
        const double4 v = dfc_[c];
        return {v.x, v.y, v.z, v.w};
    }
    __device__ __forceinline__ double geo(int comp, int c) const { return geo_[comp * (size_t)stride + c]; }
    __device__ __forceinline__ double manning(int c) const { return man_[c]; }
};

// Loader of the step kernel: the tile's own states and neighbour codes are
// staged in LDS (s_st, s_nb: cells base .. base + n - 1); a cell of the same
// tile is read from there, any other from global memory.  Same values, so
// the same bits; neighbour gathers mostly hit LDS instead of L2.
struct TileLd {
    const double4 *st_;
    const int4 *nbr_;
    const double4 *dfc_;
    const double *geo_;
    const double *man_;
    int stride;
    const double4 *s_st;
    const int4 *s_nb;
    int base, n;
    __device__ __forceinline__ St<double> state(int c) const {
        const unsigned o = (unsigned)(c - base);
        const double4 v = o < (unsigned)n ? s_st[o] : st_[c];
        return {v.x, v.y, v.z, v.w};
    }
    __device__ __forceinline__ I4 nbr(int c) const {
        const unsigned o = (unsigned)(c - base);
        const int4 v = o < (unsigned)n ? s_nb[o] : nbr_[c];
        return {v.x, v.y, v.z, v.w};
    }
    __device__ __forceinline__ D4 dfc(int c) const {
        const double4 v = dfc_[c];
        return {v.x, v.y, v.z, v.w};
    }
    __device__ __forceinline__ double geo(int comp, int c) const { return geo_[comp * (size_t)stride + c]; }
    __device__ __forceinline__ double manning(int c) const { return man_[c]; }
};

struct AdjState;  // reverse-sweep storage (cf_adjoint.hip)

}  // namespace cf

struct cf_handle {
    int ninst = 0;
    int ncell = 0;   // padded total
    int ntiles = 0;
    std::vector<int> nelt, nnt, cell_off, tile_off;    // per instance (tile_off has ninst+1)
    std::vector<cf_params> prm;
    std::vector<cf::KParams> kp;
    std::vector<std::vector<int>> hperm;               // per instance: internal -> caller (local)
    std::vector<int4> hnbr;                            // host copy of the neighbour codes
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int iflux = 1, fric = 0, skip = 1, xcd = 0;
    int tile = cf::kTile;              // cells per tile (workgroup size)
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
    double *partial = nullptr;         // per tile: cfl_kernel minima (dt of a new state)
    unsigned long long *dtslot = nullptr;  // per instance: the step's dt slots (StepArgs::dtslot)
    unsigned *cntslot = nullptr;
    unsigned *nanflag = nullptr;
    unsigned *stamp = nullptr;
    int *tile_adj_off = nullptr, *tile_adj = nullptr;
    unsigned *always = nullptr;
    int *tile_inst = nullptr, *inst_end = nullptr, *inst_tile = nullptr;
    double *surf_in = nullptr;         // ninst x CF_MAX_IO inlet areas (phys_inflow)
    int phys_in = 0;                   // an instance uses inlet = phys_inflow
    // Run record: one device block holding, in this order, a mirror of the
    // control blocks (kept current by run_init_kernel and finalize_kernel),
    // dt_hist, ran, qin_hist, qout_hist and tc_hist; a run's results come
    // back in one copy of its first part into the pinned host block hrec.
    char *rec = nullptr;
    char *hrec = nullptr;
    size_t rec_dt = 0, rec_ran = 0, rec_qin = 0, rec_qout = 0, rec_tc = 0, rec_bytes = 0;
    cf::InstCtrl *ctrl_mirror = nullptr;
    double *dt_hist = nullptr, *tc_hist = nullptr;
    int *ran = nullptr;
    int cap = 0;                       // steps per run the history arrays hold
    // discharge bookkeeping (reference check_steady_state): on when asked
    // for or when an instance has tol_reg_perm > 0
    int discharge = 0;
    int nio = 0;                       // cells with an inflow or outflow face
    int *io_cell = nullptr;            // per slot: the cell (cells with an I/O face, caller order)
    int *io_off = nullptr;             // per instance: first slot (ninst + 1)
    double *q_io = nullptr;            // 2 per slot: outflow, inflow discharge of the step
    double *qin_hist = nullptr, *qout_hist = nullptr;  // like dt_hist (run record), when discharge is on
    std::vector<std::vector<double>> last_qin, last_qout;  // per instance, last run
    // host mirror of the control block after the last finished run
    std::vector<cf::InstCtrl> hctrl;
    std::vector<long long> last_run, last_total;
    int pending = -1;
    int max_cells = 0;
    // graphs of whole runs, keyed by the step count
    struct Graph {
        int nsteps;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    int use_graphs = 1;
    // workgroup -> (tile, instance) of the step kernel; tile classes of the
    // wet-first order
    int2 *tile_map = nullptr;
    unsigned char *tile_class = nullptr;
    unsigned long long *rank_counts = nullptr;  // per 256 tiles: packed class counts of the ranking
    int wet_first = 1;                 // the wet-first order is on
    cf::AdjState *adj = nullptr;
    unsigned params_ver = 0;           // bumped by every inflow or hydrograph change
    int ad_flags = 0;                  // CF_ADJ_* bits of the derivative runs
};

// ---- host helpers shared by the translation units ----------------------

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "cuteflow_hip: %s failed: %s\n", #x, hipGetErrorString(e_));  \
            return CF_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

namespace cf {

inline int grid_of(int n, int b) { return (n + b - 1) / b; }

template <typename T>
inline void dfree(T *&ptr) {
    if (ptr) (void)hipFree((void *)ptr);
    ptr = nullptr;
}

template <typename T>
inline bool dalloc(T *&p, size_t count) {
    return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)) == hipSuccess;
}

template <typename T>
inline bool h2d(T *dst, const T *src, size_t count, hipStream_t s) {
    return count == 0 || hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s) == hipSuccess;
}

// defined in cuteflow_hip.hip
StepArgs base_args(cf_handle *h);
void launch_step(const StepArgs &a, int iflux, int fric, int grid, hipStream_t s);
void launch_finalize(cf_handle *h, double *dt_hist, double *tc_hist, int *ran, int stride, int step, hipStream_t s,
                     double *qin_hist = nullptr, double *qout_hist = nullptr, InstCtrl *mirror = nullptr);
void launch_run_init(cf_handle *h, double *dt_hist, double *tc_hist, int stride, hipStream_t s,
                     InstCtrl *mirror = nullptr);
int set_schedule(cf_handle *h, int skip, int xcd);
// the wet-first ranking of the tiles (no-op with the fixed order): by the
// classes the last step wrote, or with `ran` by `nrows` consecutive activity
// rows of a recorded run (the tiles that ran at one of those steps first);
// each class in XCD-contiguous chunks when xcd is 1 (-1: as the handle's
// schedule says)
void launch_rank(cf_handle *h, hipStream_t s, const unsigned char *ran = nullptr, int nrows = 0, int xcd = -1);
// defined in cf_adjoint.hip
void adj_free(cf_handle *h);
void adj_drop_graphs(cf_handle *h);

}  // namespace cf
