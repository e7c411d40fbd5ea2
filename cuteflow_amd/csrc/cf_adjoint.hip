// Discrete adjoint and tangent-linear model of the time step (no reference
// counterpart: the reference's Manning calibration, Calibration_manning/,
// estimates n from repeated forward runs).
//
// Recorded forward: a run of nsteps steps that keeps dt and tc of every step
// (the reverse sweep freezes them), a checkpoint of the state every K steps,
// the water level h + zm at gauge cells every `every` steps and optionally
// the per-cell maximum depth over those samples with the sample it occurred
// at.
//
// Reverse sweep, one segment of K steps at a time from the last: the
// segment is replayed from its checkpoint (same step kernel, dt/tc from the
// record, the tiles that ran in the record) keeping each step's input state
// and residual x; then for each step k, last to first, one launch of
//   revg_kernel    one thread per cell of the tiles that ran at step k:
//                  sums the contributions addressed to its cell by the
//                  reverse of step k + 1 in a fixed order (own, the
//                  neighbours' R1 slots that face it, its R2 entries by a
//                  CSR list) plus the seeds of the state after step k, then
//                  cell_reverse() writes the cotangent of the cell's own old
//                  state, of its three neighbours' (R1) and of its WAF
//                  second-ring upwind cells (R2), and adds the Manning
//                  cotangent to gn (R1/R2 double-buffered by step parity);
// and a last gather gives the initial-state cotangent.
// No atomics: every gradient is reproducible bit for bit.
//
// Tangent-linear model: the step body evaluated on dual numbers, replaying
// the recorded dt/tc, carrying the derivative of the state along a
// direction (dn, dU0); its values are the forward's bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include <type_traits>

#include "cf_dual.h"
#include "cf_dual_rec.h"
#include "cf_internal.h"

#pragma clang fp contract(off)

using namespace cf;

namespace cf {

struct AdjState {
    // configuration of the recorded run
    int nsteps = 0, every = 1, K = 1, nseg = 0, nsamp = 0, want_hmax = 0;
    bool recorded = false;
    bool tlm_ready = false;
    // host copies: the recorded gauge levels (read once per record) and the
    // gauge seeds (cf_adj_set_seed writes columns of it; uploaded once, by
    // the next reverse sweep)
    std::vector<double> eta_host, geta_host;
    bool eta_host_ok = false, geta_dirty = false;
    int gtotal = 0;
    std::vector<std::vector<int>> gcells;  // per instance: caller-local gauge cells
    std::vector<int> goff;                 // per instance: first gauge (ninst + 1)
    int *d_gcell = nullptr;                // internal cell of each gauge
    int *d_gfirst = nullptr;               // per cell: first gauge at the cell, -1 if none
    int *d_gnext = nullptr;                // per gauge: next gauge at the same cell, -1 if none
    int *d_ginst = nullptr;                // instance of each gauge
    int *d_goff = nullptr;
    double *eta = nullptr, *geta = nullptr, *deta = nullptr;  // nsamp x gtotal
    double *teta = nullptr;   // gauge levels of the tangent-linear run's values
    double *hmax = nullptr, *ghmax = nullptr, *dhmax = nullptr;
    int *argmax = nullptr;
    double *dt = nullptr, *tc = nullptr;  // ninst x (nsteps + 1)
    int *ran = nullptr;
    double4 *ckpt = nullptr;  // nseg x ncell
    double4 *seg = nullptr;   // K x ncell: states k0+1 .. k0+K of a segment
    double *xs = nullptr;     // K x 3 x ncell: residual x of each step of a segment
    double4 *lam[2] = {nullptr, nullptr};
    double4 *gstate = nullptr;  // final-state seed
    double *R1[2] = {nullptr, nullptr};   // 12 x ncell: (h, hu, hv) for neighbours 0..2, own (by step parity)
    // by step parity, one bit per cell (64 cells a word): the cell's R1/R2
    // contributions are nonzero (it was not clamped); the slots of a clamped
    // cell are neither written nor read
    unsigned long long *cmask[2] = {nullptr, nullptr};
    double4 *R2[2] = {nullptr, nullptr};  // 6 x ncell: face k, slot s at 2k + s
    int *r2_off = nullptr, *r2_idx = nullptr;
    double *gn = nullptr;
    unsigned char *tile_ran = nullptr;  // nsteps x ntiles: tile ran in the recorded run
    unsigned char *act = nullptr;       // K x ntiles: the replay's action per step of a segment and tile
    // seg and xs hold the last segment's steps as the recorded forward wrote
    // them (the reverse sweep then does not replay that segment)
    bool tape_last = false;
    int lam_cur = 0;
    // tangent-linear run
    double4 *tst[2] = {nullptr, nullptr}, *tdst[2] = {nullptr, nullptr};
    double4 *tdir = nullptr;  // tangent of the initial state
    double *dman = nullptr;   // tangent of Manning n
    int tlm_cur = 0;
    // branch masks of tangent-linear runs: mode 1 records, per step and
    // cell, whether the update left the cell wet (one bit, 64 cells a word);
    // mode 2 counts per cell the steps whose mask differs from the recorded
    // one (clamp decisions that a perturbed run takes differently); modes 3
    // and 4 also record / compare a signature of every branch of the update
    int mask_mode = 0, mask_steps = 0, sig_steps = 0;
    bool mask_rec = false, sig_rec = false;  // masks / signatures of the current record exist
    unsigned *bsig = nullptr;   // nsteps x ncell branch signatures
    int *bflips = nullptr;
    // Manning n of the recorded run (the replay and the reverse sweep use
    // it, whatever n the handle holds later) and the parameter version it
    // was recorded with (an inflow or hydrograph change since refuses the
    // reverse sweep)
    double *man_rec = nullptr;
    unsigned params_ver = 0;
    unsigned long long *wmask = nullptr;
    int *flips = nullptr;
    bool r2_built = false;
    bool gauges_dirty = true;
    int alloc_steps = 0, alloc_K = 0, alloc_nsamp = 0, alloc_gtotal = 0;
    // captured launch sequences, by kind: 0 recorded forward, 1 reverse
    // sweep, 2 TLM run, 3 reverse sweep that takes the last segment's tape
    // from the recorded forward
    struct Graph {
        int kind, nsteps, every, K, hmax;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
};

}  // namespace cf

namespace {

struct RevArgs {
    const double4 *st;
    const int4 *nbr;
    const double4 *dfc;
    const double *geo;
    const double *manning;
    const double *hydro;
    const KParams *kp;
    const int *tile_inst;
    const int *inst_end;
    const double *xs;
    const double4 *lam;        // cotangent of the state after the last step (when ran_next is null)
    const double *R1in;        // contributions of the reverse of step k + 1
    const double4 *R2in;
    const unsigned long long *cmask_in;  // cells of step k + 1 with nonzero contributions
    unsigned long long *cmask;           // the same for step k
    double *R1;                // contributions of this step's reverse
    double4 *R2;
    const int *r2_off;
    const int *r2_idx;
    double *gn;
    const double *dt;
    const double *tc;
    const unsigned char *ran;       // this step's row of tile activity
    const unsigned char *ran_next;  // step k + 1's row; null for the last step
    // seeds of the state after this step: hmax of sample sidx, gauge
    // levels of that sample (geta_row, null if none; gauges of a cell
    // listed by gfirst/gnext in increasing index)
    const int *argmax;
    const double *ghmax;
    const double *geta_row;
    const int *gfirst;
    const int *gnext;
    int sidx;
    int tile_cells;
    int hstride;
    int step;
    int ncell;
    int xcd;  // workgroup -> tile order (tile_of_block) when tile_map is null
    // workgroup -> (tile, instance): the forward's wet-first map, so that the
    // tiles that ran (wet ones first) take the first workgroups
    const int2 *tile_map;
};

// Occupancy target of the reverse kernel (waves per SIMD): the first-order
// reverse fits 168 VGPRs (3 waves) with a few spills and runs 8 % faster
// than at 2 waves on MI355X; the WAF and MUSCL reverses need 256 (1 wave).
#ifndef CF_REV_WAVES_FIRST_ORDER
#define CF_REV_WAVES_FIRST_ORDER 3
#endif
template <int IFLUX>
struct RevWaves {
    static constexpr int value = IFLUX >= 3 ? 1 : CF_REV_WAVES_FIRST_ORDER;
};

__device__ __forceinline__ void acc4(double4 &a, const double4 &b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
}

// Cotangent of cell c's state after step k: the contributions the reverse
// of step k + 1 addressed to it, summed in a fixed order (own, the
// neighbours' R1 slots that face it, its R2 entries by a CSR list;
// contributions of tiles that did not run at step k + 1 are zero and not
// read), then the seeds of that state (hmax of sample sidx, then the gauge
// levels of that sample in increasing gauge index).  `tiles_ran` is the
// activity row of step k + 1; a cell whose tile did not run gets its seeds
// only.
// Contribution slots: per cell 12 doubles, (h, hu, hv) for the cells
// across faces 0..2, then its own.
constexpr int kR1 = 12;

// tiles are 64, 128 or 256 cells: a cell's tile by a shift (tshift = log2 of
// the tile size) rather than an integer division by a runtime value
__device__ __forceinline__ bool live_cell(const unsigned long long *__restrict__ cm, int c) {
    return ((cm[c >> 6] >> (c & 63)) & 1ull) != 0ull;
}

// `cm`: the cells of step k + 1 whose contributions are nonzero; the slots
// of the others are zero and not added (skipping the addition of +0 changes
// at most the sign of a zero sum).  Every load whose address is known is
// issued before any test of a loaded value: the codes, the activity bytes and
// mask words of the cell and its neighbours and all four slots (a missing
// neighbour reads a slot of the cell's own row, unused), so that the gather
// takes two rounds of loads instead of a chain of a dozen.
template <bool R2ON>
__device__ __forceinline__ double4 gather_cell(int c, int tshift, const unsigned char *__restrict__ tiles_ran,
                                               const unsigned long long *__restrict__ cm,
                                               const int4 *__restrict__ nbr, const double *__restrict__ R1,
                                               const double4 *__restrict__ R2, const int *__restrict__ r2_off,
                                               const int *__restrict__ r2_idx, const int *__restrict__ argmax,
                                               const double *__restrict__ ghmax, int sidx,
                                               const double *__restrict__ geta_row, const int *__restrict__ gfirst,
                                               const int *__restrict__ gnext) {
    const int4 nb = nbr[c];
    const int nbv[3] = {nb.x, nb.y, nb.z};
    const int am = argmax ? argmax[c] : -1;
    const double gh = argmax ? ghmax[c] : 0.0;
    const int g0 = geta_row ? gfirst[c] : -1;
    const int ran_c = tiles_ran[c >> tshift];
    const unsigned long long wc = cm[c >> 6];
    const double *o = R1 + kR1 * (size_t)c + 9;
    const double o0 = o[0], o1 = o[1], o2 = o[2];
    int ii[3], ran_k[3];
    unsigned long long w_k[3];
    double v[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        ii[k] = nbv[k] >= 0 ? nbv[k] : c;
        ran_k[k] = tiles_ran[ii[k] >> tshift];
        w_k[k] = cm[ii[k] >> 6];
        const double *sk = R1 + kR1 * (size_t)ii[k] + 3 * ((nb.w >> (2 * k)) & 3);
        v[k][0] = sk[0];
        v[k][1] = sk[1];
        v[k][2] = sk[2];
    }
    double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
    if (ran_c) {
        if ((wc >> (c & 63)) & 1ull) acc = make_double4(o0, o1, o2, 0.0);
#pragma unroll
        for (int k = 0; k < 3; k++)
            if ((nbv[k] >= 0) & (ran_k[k] != 0) & (((w_k[k] >> (ii[k] & 63)) & 1ull) != 0ull)) {
                acc.x += v[k][0];
                acc.y += v[k][1];
                acc.z += v[k][2];
            }
        if (R2ON)
            for (int e = r2_off[c]; e < r2_off[c + 1]; e++) {
                const int src = r2_idx[e], sc = src / 6;
                if (tiles_ran[sc >> tshift] && live_cell(cm, sc)) acc4(acc, R2[src]);
            }
    }
    if (argmax && am == sidx) acc.x += gh;
    for (int g = g0; g >= 0; g = gnext[g]) acc.x += geta_row[g];
    acc.w = 0.0;
    return acc;
}

// Reverse of step k, one thread per cell of the tiles that ran at step k
// (a tile skipped in the recorded run holds dry cells with dry neighbours:
// every contribution of its reverse is zero).  The thread first forms its
// cell's cotangent of the state after step k from the contributions of
// step k + 1's reverse (gather_cell; for the last step it is read from
// r.lam), then cell_reverse() gives the cotangent of the cell's own old
// state, of its three neighbours' (R1) and of its WAF second-ring upwind
// cells (R2), and the Manning cotangent added to gn.  R1/R2 alternate
// between two buffers from step to step, so the gather of step k + 1 and
// the reverse of step k share one launch.
template <int IFLUX, int FRIC>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(RevWaves<IFLUX>::value, 8))) void revg_kernel(
    RevArgs r, const KParams *__restrict__ kp) {
    int tile, b;
    if (r.tile_map) {
        const int2 tb = r.tile_map[blockIdx.x];
        tile = tb.x;
        b = tb.y;
    } else {
        tile = tile_of_block(r.xcd);
        b = r.tile_inst[tile];
    }
    const int i = tile * (int)blockDim.x + threadIdx.x;
    if (i >= r.inst_end[b] || !r.ran[tile]) return;
    const size_t hix = (size_t)b * r.hstride + r.step;
    const double x[3] = {r.xs[i], r.xs[(size_t)r.ncell + i], r.xs[2 * (size_t)r.ncell + i]};
    // A cell whose new state the forward clamped to dry contributes nothing
    // whatever its cotangent (cell_reverse returns zeros at the same test):
    // its cotangent is not gathered, its slots are not written (its bit in
    // r.cmask tells the gather of step k - 1 not to read them) and its
    // Manning cotangent is +0 (gn, a sum from +0, is never -0: unchanged).
    const double r0 = (FRIC == 0 || FRIC == 1) ? x[0] : (1.0 * x[0] + 0.0 * x[1]) + 0.0 * x[2];
    const bool clamped = r.st[i].x + r.dt[hix] * r0 <= kp[b].tolisec;
    // the real cells of a tile come first: lane 0 of a wave with any is one
    const unsigned long long live = __ballot(!clamped);
    if ((threadIdx.x & 63) == 0) r.cmask[i >> 6] = live;
    if (clamped) return;
    const double4 l4 = r.ran_next ? gather_cell<IFLUX >= 3>(i, __ffs((int)blockDim.x) - 1, r.ran_next, r.cmask_in,
                                                            r.nbr, r.R1in, r.R2in, r.r2_off, r.r2_idx, r.argmax,
                                                            r.ghmax, r.sidx, r.geta_row, r.gfirst, r.gnext)
                                  : r.lam[i];
    const DevLd ld{r.st, r.nbr, r.dfc, r.geo, r.manning, r.ncell};
    const double lam[3] = {l4.x, l4.y, l4.z};
    double own[3], r1[3][3], r2[3][2][3], gn;
    cell_reverse<IFLUX, FRIC>(ld, kp[b], r.hydro, i, r.dt[hix], r.tc[hix], x, lam, own, r1, r2, gn);
    double *slots = r.R1 + kR1 * (size_t)i;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        slots[3 * k] = r1[k][0];
        slots[3 * k + 1] = r1[k][1];
        slots[3 * k + 2] = r1[k][2];
    }
    slots[9] = own[0];
    slots[10] = own[1];
    slots[11] = own[2];
    if (IFLUX >= 3) {
        const int mask = r.nbr[i].w >> kR2Shift;
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int s = 0; s < 2; s++)
                if (mask & (1 << (2 * k + s)))
                    r.R2[6 * (size_t)i + 2 * k + s] = make_double4(r2[k][s][0], r2[k][s][1], r2[k][s][2], 0.0);
    }
    r.gn[i] += gn;
}

// Cotangent of the initial state: the gather of step 0's contributions
// (the initial state carries no seeds), every cell.
template <bool R2ON>
__global__ void gather_kernel(const int4 *__restrict__ nbr, const double *__restrict__ R1,
                              const double4 *__restrict__ R2, const int *__restrict__ r2_off,
                              const int *__restrict__ r2_idx, const int *__restrict__ tile_inst,
                              const int *__restrict__ inst_end, const unsigned char *__restrict__ ran,
                              const unsigned long long *__restrict__ cm, double4 *__restrict__ lam_out) {
    const int tile = blockIdx.x;
    const int b = tile_inst[tile];
    const int c = tile * (int)blockDim.x + threadIdx.x;
    if (c >= inst_end[b]) return;
    lam_out[c] = gather_cell<R2ON>(c, __ffs((int)blockDim.x) - 1, ran, cm, nbr, R1, R2, r2_off, r2_idx, nullptr, nullptr,
                                   -1, nullptr, nullptr, nullptr);
}

// lam = final-state seed + hmax seed of the last sample (cells not seeded: 0)
__global__ void lam_init_kernel(const int *__restrict__ tile_inst, const int *__restrict__ inst_end,
                                const double4 *__restrict__ gstate, const int *__restrict__ argmax,
                                const double *__restrict__ ghmax, int sidx, double4 *__restrict__ lam) {
    const int tile = blockIdx.x;
    const int b = tile_inst[tile];
    const int c = tile * (int)blockDim.x + threadIdx.x;
    double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
    if (c < inst_end[b]) {
        v = gstate[c];
        v.w = 0.0;
        if (argmax && argmax[c] == sidx) v.x += ghmax[c];
    }
    lam[c] = v;
}

// Gauge seeds of sample sidx, one thread per instance (gauges of an
// instance added in order).
__global__ void gauge_seed_kernel(const int *__restrict__ gcell, const int *__restrict__ goff,
                                  const double *__restrict__ geta, int gtotal, int sidx, int ninst,
                                  double4 *__restrict__ lam) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ninst) return;
    for (int g = goff[b]; g < goff[b + 1]; g++) lam[gcell[g]].x += geta[(size_t)sidx * gtotal + g];
}

// Current state buffer of the instance of `tile` during a run.
__device__ __forceinline__ const double4 *cur_state(double4 *const *st, const InstCtrl *ctrl, int b) {
    const InstCtrl &c = ctrl[b];
    return st[(c.par + c.steps) & 1];
}

struct StPair {
    double4 *s[2];
};

__global__ void sample_gauges_kernel(StPair st, const InstCtrl *ctrl, const int *__restrict__ gcell,
                                     const int *__restrict__ ginst, int gtotal, int sidx, double *__restrict__ eta) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= gtotal) return;
    const double4 u = cur_state(st.s, ctrl, ginst[g])[gcell[g]];
    eta[(size_t)sidx * gtotal + g] = u.x + u.w;
}

// Maximum depth over the samples (sample sidx, after step k).  `ran`
// holds the activity rows of the `every` steps since the previous sample:
// a tile that ran at none of them holds that sample's state bit for bit (a
// skipped tile writes nothing and both state buffers hold its dry state), so
// its maxima stand and the tile is not read.  Several tiles per workgroup.
__global__ void sample_hmax_kernel(StPair st, const InstCtrl *ctrl, const int *__restrict__ tile_inst,
                                   const int *__restrict__ inst_end, const unsigned char *__restrict__ ran,
                                   int ntiles, int every, int tshift, int sidx, double *__restrict__ hmax,
                                   int *__restrict__ argmax) {
    const int c = blockIdx.x * (int)blockDim.x + threadIdx.x;
    // tiles are 64 cells or more: one tile per wave
    const int tile = __builtin_amdgcn_readfirstlane(c >> tshift);
    if (tile >= ntiles) return;
    if (sidx > 0) {
        // lane j reads the row of step j: one round of loads, not a chain
        const int lane = (int)(threadIdx.x & 63);
        bool any = false;
        for (int j0 = 0; j0 < every && !any; j0 += 64) {
            const int j = j0 + lane;
            any = __any(j < every && ran[(size_t)j * ntiles + tile] != 0);
        }
        if (!any) return;
    }
    const int b = tile_inst[tile];
    if (c >= inst_end[b]) return;
    const double h = cur_state(st.s, ctrl, b)[c].x;
    if (sidx == 0 || h > hmax[c]) {
        hmax[c] = h;
        argmax[c] = sidx;
    }
}

// The replay's action for each tile at steps k0 .. k0 + n - 1 of a segment
// (row m): run the tile if it ran in the recorded run; else rewrite its
// output slot to the dry state unless the slot is known to hold it already
// (the slot last held step k0 + m + K, rows `next`, for m < nnext, and the
// tile was skipped then too).
__global__ void replay_action_kernel(const unsigned char *__restrict__ ran, int n, int ntiles,
                                     const unsigned char *__restrict__ next, int nnext,
                                     unsigned char *__restrict__ act) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    for (int m = 0; m < n; m++) {
        const size_t o = (size_t)m * ntiles + t;
        act[o] = ran[o] ? kReplayRun : ((next && m < nnext && !next[o]) ? kReplayKeep : kReplayRewrite);
    }
}

__global__ void ckpt_kernel(StPair st, const InstCtrl *ctrl, const int *__restrict__ tile_inst,
                            const int *__restrict__ inst_end, double4 *__restrict__ dst) {
    const int tile = blockIdx.x;
    const int b = tile_inst[tile];
    const int c = tile * (int)blockDim.x + threadIdx.x;
    if (c >= inst_end[b]) return;
    dst[c] = cur_state(st.s, ctrl, b)[c];
}

// ---- tangent-linear model ------------------------------------------------

template <class D>
struct TlmLd {
    const double4 *st_, *dst_;
    const int4 *nbr_;
    const double4 *dfc_;
    const double *geo_, *man_, *dman_;
    int stride;
    __device__ __forceinline__ St<D> state(int c) const {
        const double4 v = st_[c], d = dst_[c];
        return {D(v.x, d.x), D(v.y, d.y), D(v.z, d.z), v.w};
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
    __device__ __forceinline__ D manning(int c) const { return D(man_[c], dman_[c]); }
};

struct TlmArgs {
    const double4 *st_in, *dst_in;
    double4 *st_out, *dst_out;
    const int4 *nbr;
    const double4 *dfc;
    const double *geo;
    const double *manning, *dman;
    const double *hydro;
    const KParams *kp;
    const int *tile_inst, *inst_end;
    const double *dt, *tc;
    int hstride, step, ncell, tile_cells, xcd;
    const int2 *tile_map;  // workgroup -> (tile, instance), as in the forward (null: tile_of_block)
    // Diagnostic masks (0 none): 1 / 3 record, 2 / 4 count the differences
    // from the recorded run.  Every mode > 0 keeps the wet bit of each (step,
    // cell); modes 3 and 4 also a signature of every branch the cell's update
    // took (a REC kernel).
    int mask_mode;
    unsigned long long *wmask;     // nsteps x (ncell / 64) words
    int *flips;                    // per cell: steps whose wet bit differed
    unsigned *bsig;                // nsteps x ncell branch signatures
    int *bflips;                   // per cell: steps whose signature differed
};

template <int IFLUX, int FRIC, bool REC>
__global__ __launch_bounds__(kTile) void tlm_kernel(TlmArgs a) {
    using D = typename std::conditional<REC, DualR, Dual>::type;
    int tile, b;
    if (a.tile_map) {
        const int2 tb = a.tile_map[blockIdx.x];
        tile = tb.x;
        b = tb.y;
    } else {
        tile = tile_of_block(a.xcd);
        b = a.tile_inst[tile];
    }
    const int i = tile * (int)blockDim.x + threadIdx.x;
    const bool valid = i < a.inst_end[b];
    double h_new = 0.0;
    unsigned sig = 0u;
    if (valid) {
        if (REC) s_branch_sig[threadIdx.x] = 0x811c9dc5u;
        const TlmLd<D> ld{a.st_in, a.dst_in, a.nbr, a.dfc, a.geo, a.manning, a.dman, a.ncell};
        const size_t hix = (size_t)b * a.hstride + a.step;
        D o[3];
        cell_update<D, IFLUX, FRIC>(ld, a.kp[b], a.hydro, i, a.dt[hix], a.tc[hix], o, (D *)nullptr);
        const double zm = a.st_in[i].w;
        a.st_out[i] = make_double4(o[0].v, o[1].v, o[2].v, zm);
        a.dst_out[i] = make_double4(o[0].d, o[1].d, o[2].d, 0.0);
        h_new = o[0].v;
        if (REC) sig = s_branch_sig[threadIdx.x];
    }
    if (a.mask_mode) {
        const bool record = a.mask_mode == 1 || a.mask_mode == 3;
        // the clamp sets h to +0 exactly and an unclamped depth exceeds the
        // dry threshold: h > 0 is "the clamp was not taken"
        const bool wet = valid && h_new > 0.0;
        const unsigned long long bal = __ballot(wet);
        const size_t w = (size_t)a.step * (size_t)(a.ncell >> 6) + (size_t)(i >> 6);
        if (record) {
            if ((threadIdx.x & 63) == 0) a.wmask[w] = bal;
        } else if (valid) {
            const bool was = ((a.wmask[w] >> (i & 63)) & 1ull) != 0ull;
            if (was != wet) a.flips[i] += 1;
        }
        if (REC && valid) {
            const size_t ws = (size_t)a.step * (size_t)a.ncell + (size_t)i;
            if (record) a.bsig[ws] = sig;
            else if (a.bsig[ws] != sig) a.bflips[i] += 1;
        }
    }
}

__global__ void int_out_kernel(const int *__restrict__ src, const int *__restrict__ perm, int c0, int n,
                               int *__restrict__ dst) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) dst[perm[c0 + l]] = src[c0 + l];
}

__global__ void fill_i_kernel(int *p, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void tlm_sample_kernel(const double4 *__restrict__ st, const double4 *__restrict__ dst,
                                  const int *__restrict__ gcell, int gtotal, int sidx, double *__restrict__ teta,
                                  double *__restrict__ deta) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= gtotal) return;
    const double4 u = st[gcell[g]];
    teta[(size_t)sidx * gtotal + g] = u.x + u.w;
    deta[(size_t)sidx * gtotal + g] = dst[gcell[g]].x;
}

__global__ void tlm_hmax_kernel(const double4 *__restrict__ dst, const int *__restrict__ tile_inst,
                                const int *__restrict__ inst_end, const int *__restrict__ argmax, int sidx,
                                double *__restrict__ dhmax) {
    const int tile = blockIdx.x;
    const int b = tile_inst[tile];
    const int c = tile * (int)blockDim.x + threadIdx.x;
    if (c >= inst_end[b]) return;
    if (argmax[c] == sidx) dhmax[c] = dst[c].x;
}

// ---- small helpers -------------------------------------------------------

__global__ void fill_d4_kernel(double4 *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

__global__ void fill_d_kernel(double *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void copy_d4_kernel(const double4 *__restrict__ src, double4 *__restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// caller order (local index perm[c]) -> internal cell c, per-cell scalar
__global__ void scalar_in_kernel(const double *__restrict__ src, const int *__restrict__ perm, int c0, int n,
                                 double *__restrict__ dst) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) dst[c0 + l] = src[perm[c0 + l]];
}

__global__ void scalar_out_kernel(const double *__restrict__ src, const int *__restrict__ perm, int c0, int n,
                                  double *__restrict__ dst) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) dst[perm[c0 + l]] = src[c0 + l];
}

// (h, hu, hv) per cell in caller order <-> double4 (zm slot left 0)
__global__ void state3_in_kernel(const double *__restrict__ src, const int *__restrict__ perm, int c0, int n,
                                 double4 *__restrict__ dst) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) {
        const size_t o = 3 * (size_t)perm[c0 + l];
        dst[c0 + l] = make_double4(src[o], src[o + 1], src[o + 2], 0.0);
    }
}

__global__ void state3_out_kernel(const double4 *__restrict__ src, const int *__restrict__ perm, int c0, int n,
                                  double *__restrict__ dst) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) {
        const double4 v = src[c0 + l];
        const size_t o = 3 * (size_t)perm[c0 + l];
        dst[o] = v.x;
        dst[o + 1] = v.y;
        dst[o + 2] = v.z;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

static void fill0(double4 *p, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fill_d4_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}
static void fill0(double *p, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fill_d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

static void drop_adj_graphs(AdjState *A) {
    for (auto &g : A->graphs) (void)hipGraphExecDestroy(g.exec);
    A->graphs.clear();
}

void cf::adj_drop_graphs(cf_handle *h) {
    if (h && h->adj) drop_adj_graphs(h->adj);
}

void cf::adj_free(cf_handle *h) {
    AdjState *A = h ? h->adj : nullptr;
    if (!A) return;
    drop_adj_graphs(A);
    dfree(A->d_gcell); dfree(A->d_ginst); dfree(A->d_goff); dfree(A->d_gfirst); dfree(A->d_gnext);
    dfree(A->eta); dfree(A->geta); dfree(A->deta); dfree(A->teta);
    dfree(A->hmax); dfree(A->ghmax); dfree(A->dhmax); dfree(A->argmax);
    dfree(A->dt); dfree(A->tc); dfree(A->ran);
    dfree(A->ckpt); dfree(A->seg); dfree(A->xs);
    dfree(A->lam[0]); dfree(A->lam[1]); dfree(A->gstate);
    dfree(A->R1[0]); dfree(A->R1[1]); dfree(A->cmask[0]); dfree(A->cmask[1]); dfree(A->R2[0]); dfree(A->R2[1]); dfree(A->r2_off); dfree(A->r2_idx); dfree(A->gn); dfree(A->tile_ran); dfree(A->act);
    dfree(A->tst[0]); dfree(A->tst[1]); dfree(A->tdst[0]); dfree(A->tdst[1]); dfree(A->tdir); dfree(A->dman);
    dfree(A->wmask); dfree(A->flips); dfree(A->man_rec); dfree(A->bsig); dfree(A->bflips);
    delete A;
    h->adj = nullptr;
}

static AdjState *adj_of(cf_handle *h) {
    if (!h->adj) {
        h->adj = new (std::nothrow) AdjState();
        if (h->adj) {
            h->adj->gcells.resize(h->ninst);
            h->adj->goff.assign(h->ninst + 1, 0);
        }
    }
    return h->adj;
}

// Second-ring adjoint targets of the WAF stencil: marks in nbr.w which
// (face, slot) of each cell addresses a cell through R2 and builds the
// inverse lists (target cell -> R2 entries), both from the topology alone
// with the same rules as the forward's upwind choice.
static int build_r2(cf_handle *h, AdjState *A) {
    if (A->r2_built) return CF_OK;
    const int N = h->ncell;
    std::vector<int> tgt(6 * (size_t)N, -1);
    std::vector<int4> &nb = h->hnbr;
    for (int b = 0; b < h->ninst; b++) {
        const int c0 = h->cell_off[b], c1 = c0 + h->nelt[b];
        const int outlet = h->prm[b].outlet;
        for (int i = c0; i < c1; i++) {
            const int4 v = nb[i];
            const int c[3] = {v.x, v.y, v.z};
            int mask = 0;
            for (int j = 0; j < 3; j++) {
                int kv = -1;
                if (c[j] >= 0) kv = c[j];
                else if (((-c[j]) & 3) == 2 && outlet == 2) kv = i;
                if (kv < 0) continue;
                const int ja = (j + 1) % 3, jb = (j + 2) % 3;
                const bool own1 = c[ja] < 0, own2 = c[jb] < 0;
                const int tjv = (v.w >> (2 * j)) & 3;
                const int4 vk = (kv == i) ? v : nb[kv];
                const int ck[3] = {vk.x, vk.y, vk.z};
                const int e1 = ck[(tjv + 1) % 3], e2 = ck[(tjv + 2) % 3];
                const int kw1 = e1 >= 0 ? e1 : kv, kw2 = e2 >= 0 ? e2 : kv;
                if ((own1 && own2) || (kw1 == kv && kw2 == kv)) continue;
                if (kw1 != kv) {
                    mask |= 1 << (2 * j);
                    tgt[6 * (size_t)i + 2 * j] = kw1;
                }
                if (kw2 != kv) {
                    mask |= 1 << (2 * j + 1);
                    tgt[6 * (size_t)i + 2 * j + 1] = kw2;
                }
            }
            nb[i].w = (v.w & ((1 << kR2Shift) - 1)) | (mask << kR2Shift);
        }
    }
    std::vector<int> off(N + 1, 0);
    for (size_t e = 0; e < tgt.size(); e++)
        if (tgt[e] >= 0) off[tgt[e] + 1]++;
    for (int c = 0; c < N; c++) off[c + 1] += off[c];
    std::vector<int> idx(off[N]), fillp(off.begin(), off.end() - 1);
    for (size_t e = 0; e < tgt.size(); e++)
        if (tgt[e] >= 0) idx[fillp[tgt[e]]++] = (int)e;
    if (!dalloc(A->r2_off, off.size()) || !dalloc(A->r2_idx, idx.size()) || !dalloc(A->R2[0], 6 * (size_t)N) ||
        !dalloc(A->R2[1], 6 * (size_t)N))
        return CF_ERR_NOMEM;
    if (!h2d(A->r2_off, off.data(), off.size(), h->stream) || !h2d(A->r2_idx, idx.data(), idx.size(), h->stream) ||
        !h2d(h->nbr, nb.data(), nb.size(), h->stream))
        return CF_ERR_HIP;
    fill0(A->R2[0], 6 * (size_t)N, h->stream);
    fill0(A->R2[1], 6 * (size_t)N, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
    A->r2_built = true;
    drop_adj_graphs(A);
    return CF_OK;
}

extern "C" int cf_adj_set_flags(cf_handle *h, int flags) {
    if (!h || h->pending >= 0 || (flags & ~CF_ADJ_FREEZE_LIMITERS)) return CF_ERR_ARG;
    h->ad_flags = flags;
    for (int b = 0; b < h->ninst; b++) h->kp[b].ad_flags = flags;
    HIPCHK(hipMemcpyAsync(h->d_kp, h->kp.data(), h->kp.size() * sizeof(KParams), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->adj) h->adj->tlm_ready = false;
    return CF_OK;
}

extern "C" int cf_adj_set_gauges(cf_handle *h, int inst, int n, const int *cells) {
    if (!h || inst < 0 || inst >= h->ninst || n < 0 || (n > 0 && !cells) || h->pending >= 0) return CF_ERR_ARG;
    AdjState *A = adj_of(h);
    if (!A) return CF_ERR_NOMEM;
    for (int g = 0; g < n; g++)
        if (cells[g] < 0 || cells[g] >= h->nelt[inst]) return CF_ERR_ARG;
    A->gcells[inst].assign(cells, cells + n);
    A->gauges_dirty = true;
    A->recorded = false;
    A->tlm_ready = false;
    return CF_OK;
}

static int upload_gauges(cf_handle *h, AdjState *A) {
    if (!A->gauges_dirty) return CF_OK;
    drop_adj_graphs(A);
    A->gauges_dirty = false;
    const int B = h->ninst;
    A->goff.assign(B + 1, 0);
    for (int b = 0; b < B; b++) A->goff[b + 1] = A->goff[b] + (int)A->gcells[b].size();
    A->gtotal = A->goff[B];
    std::vector<int> gc(std::max(A->gtotal, 1)), gi(std::max(A->gtotal, 1));
    for (int b = 0; b < B; b++) {
        std::vector<int> inv(h->nelt[b]);
        for (int l = 0; l < h->nelt[b]; l++) inv[h->hperm[b][l]] = l;
        for (size_t g = 0; g < A->gcells[b].size(); g++) {
            gc[A->goff[b] + g] = h->cell_off[b] + inv[A->gcells[b][g]];
            gi[A->goff[b] + g] = b;
        }
    }
    // gauges of each cell as a list in increasing gauge index
    std::vector<int> gfirst(h->ncell, -1), gnext(std::max(A->gtotal, 1), -1);
    for (int g = A->gtotal - 1; g >= 0; g--) {
        gnext[g] = gfirst[gc[g]];
        gfirst[gc[g]] = g;
    }
    dfree(A->d_gcell); dfree(A->d_ginst); dfree(A->d_goff); dfree(A->d_gfirst); dfree(A->d_gnext);
    if (!dalloc(A->d_gcell, gc.size()) || !dalloc(A->d_ginst, gi.size()) || !dalloc(A->d_goff, (size_t)B + 1) ||
        !dalloc(A->d_gfirst, gfirst.size()) || !dalloc(A->d_gnext, gnext.size()))
        return CF_ERR_NOMEM;
    if (!h2d(A->d_gcell, gc.data(), gc.size(), h->stream) || !h2d(A->d_ginst, gi.data(), gi.size(), h->stream) ||
        !h2d(A->d_goff, A->goff.data(), (size_t)B + 1, h->stream) ||
        !h2d(A->d_gfirst, gfirst.data(), gfirst.size(), h->stream) ||
        !h2d(A->d_gnext, gnext.data(), gnext.size(), h->stream))
        return CF_ERR_HIP;
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

// (Re)allocates the storage of a recorded run of nsteps steps.
static int adj_alloc(cf_handle *h, AdjState *A) {
    const size_t N = (size_t)h->ncell, B = (size_t)h->ninst;
    const size_t S = (size_t)A->nsteps + 1;
    if (A->alloc_steps != A->nsteps || A->alloc_K != A->K) {
        drop_adj_graphs(A);
        dfree(A->dt); dfree(A->tc); dfree(A->ran); dfree(A->ckpt); dfree(A->seg); dfree(A->xs); dfree(A->tile_ran);
        dfree(A->act);
        if (!dalloc(A->dt, B * S) || !dalloc(A->tc, B * S) || !dalloc(A->ran, B * S) ||
            !dalloc(A->ckpt, (size_t)A->nseg * N) || !dalloc(A->seg, (size_t)A->K * N) ||
            !dalloc(A->xs, (size_t)A->K * 3 * N) || !dalloc(A->tile_ran, (size_t)A->nsteps * h->ntiles) ||
            !dalloc(A->act, (size_t)A->K * h->ntiles))
            return CF_ERR_NOMEM;
        A->alloc_steps = A->nsteps;
        A->alloc_K = A->K;
    }
    if (A->alloc_nsamp != A->nsamp || A->alloc_gtotal != A->gtotal) {
        drop_adj_graphs(A);
        dfree(A->eta); dfree(A->geta); dfree(A->deta); dfree(A->teta);
        const size_t ne = (size_t)std::max(A->nsamp, 1) * (size_t)std::max(A->gtotal, 1);
        if (!dalloc(A->eta, ne) || !dalloc(A->geta, ne) || !dalloc(A->deta, ne) || !dalloc(A->teta, ne))
            return CF_ERR_NOMEM;
        A->alloc_nsamp = A->nsamp;
        A->alloc_gtotal = A->gtotal;
    }
    if (!A->lam[0]) {
        drop_adj_graphs(A);
        if (!dalloc(A->lam[0], N) || !dalloc(A->man_rec, N) || !dalloc(A->lam[1], N) || !dalloc(A->gstate, N) || !dalloc(A->R1[0], kR1 * N) ||
            !dalloc(A->R1[1], kR1 * N) || !dalloc(A->gn, N) || !dalloc(A->hmax, N) || !dalloc(A->ghmax, N) ||
            !dalloc(A->dhmax, N) || !dalloc(A->argmax, N) || !dalloc(A->cmask[0], N / 64 + 1) ||
            !dalloc(A->cmask[1], N / 64 + 1))
            return CF_ERR_NOMEM;
        fill0(A->R1[0], kR1 * N, h->stream);
        fill0(A->R1[1], kR1 * N, h->stream);
    }
    return CF_OK;
}

// Runs a launch sequence of the current recorded-run configuration: captured
// in a hipGraph on first use and replayed afterwards (the sequences' kernel
// arguments are fixed buffers; per-run data lives in device memory).
template <class F>
static int run_captured(cf_handle *h, AdjState *A, int kind, F enqueue) {
    hipStream_t s = h->stream;
    if (!h->use_graphs || !s) {
        enqueue(h, A, s);
        HIPCHK(hipGetLastError());
        return CF_OK;
    }
    for (auto &g : A->graphs)
        if (g.kind == kind && g.nsteps == A->nsteps && g.every == A->every && g.K == A->K &&
            g.hmax == A->want_hmax) {
            HIPCHK(hipGraphLaunch(g.exec, s));
            return CF_OK;
        }
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue(h, A, s);
    HIPCHK(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    A->graphs.push_back({kind, A->nsteps, A->every, A->K, A->want_hmax, exec});
    HIPCHK(hipGraphLaunch(exec, s));
    return CF_OK;
}

// Recorded forward: steps with the tile activity, samples and checkpoints.
static void enqueue_forward(cf_handle *h, AdjState *A, hipStream_t s) {
    const int S = A->nsteps + 1;
    const StPair sp{{h->st[0], h->st[1]}};
    launch_run_init(h, A->dt, A->tc, S, s);
    hipLaunchKernelGGL(ckpt_kernel, dim3(h->ntiles), dim3(h->tile), 0, s, sp, h->ctrl, h->tile_inst, h->inst_end,
                       A->ckpt);
    StepArgs a = base_args(h);
    a.dt_hist = A->dt;
    a.tc_hist = A->tc;
    a.ran = A->ran;
    a.tile_ran = A->tile_ran;
    a.hist_stride = S;
    // the steps of the last segment also write the tape the reverse sweep
    // reads (states to seg, residuals to xs), so that it replays one segment
    // less
    const int klast = (A->nseg - 1) * A->K;
    for (int k = 0; k < A->nsteps; k++) {
        a.step = k;
        if (k >= klast) {
            a.out_x = A->seg + (size_t)(k - klast) * h->ncell;
            a.xs = A->xs + (size_t)(k - klast) * 3 * h->ncell;
        }
        if (k % kReorderEvery == 0) launch_rank(h, s);
        launch_step(a, h->iflux, h->fric, h->ntiles, s);
        launch_finalize(h, A->dt, A->tc, A->ran, S, k, s);
        if ((k + 1) % A->every == 0) {
            const int sidx = (k + 1) / A->every - 1;
            if (A->gtotal)
                hipLaunchKernelGGL(sample_gauges_kernel, dim3(grid_of(A->gtotal, 256)), dim3(256), 0, s, sp, h->ctrl,
                                   A->d_gcell, A->d_ginst, A->gtotal, sidx, A->eta);
            if (A->want_hmax)
                hipLaunchKernelGGL(sample_hmax_kernel, dim3(grid_of(h->ntiles * h->tile, 256)), dim3(256), 0, s, sp,
                                   h->ctrl, h->tile_inst, h->inst_end,
                                   A->tile_ran + (size_t)(k + 1 - A->every) * h->ntiles, h->ntiles, A->every,
                                   __builtin_ctz((unsigned)h->tile), sidx, A->hmax, A->argmax);
        }
        if ((k + 1) % A->K == 0 && k + 1 < A->nsteps)
            hipLaunchKernelGGL(ckpt_kernel, dim3(h->ntiles), dim3(h->tile), 0, s, sp, h->ctrl, h->tile_inst,
                               h->inst_end, A->ckpt + (size_t)((k + 1) / A->K) * h->ncell);
    }
}

extern "C" int cf_adj_forward(cf_handle *h, int nsteps, int sample_every, int checkpoint_every, int want_hmax) {
    if (!h || nsteps <= 0 || sample_every <= 0 || checkpoint_every <= 0 || h->pending >= 0) return CF_ERR_ARG;
    // phys_inflow couples every inlet face to the inlet's summed depth: its
    // derivative is not provided
    for (int b = 0; b < h->ninst; b++)
        if (h->hctrl[b].status != 0 || h->prm[b].is_override_manning || h->prm[b].inlet == 3) return CF_ERR_ARG;
    AdjState *A = adj_of(h);
    if (!A) return CF_ERR_NOMEM;
    A->recorded = false;
    A->tlm_ready = false;
    A->tape_last = false;
    A->eta_host_ok = false;
    if (A->mask_mode) drop_adj_graphs(A);
    A->mask_mode = 0;
    A->mask_rec = A->sig_rec = false;
    A->nsteps = nsteps;
    A->every = sample_every;
    A->K = std::min(checkpoint_every, nsteps);
    A->nseg = (nsteps + A->K - 1) / A->K;
    A->nsamp = nsteps / sample_every;
    A->want_hmax = want_hmax ? 1 : 0;
    int rc = upload_gauges(h, A);
    if (rc) return rc;
    if ((rc = adj_alloc(h, A)) != CF_OK) return rc;
    if (h->iflux >= 3 && (rc = build_r2(h, A)) != CF_OK) return rc;
    // seeds default to zero for instances the caller does not seed
    fill0(A->geta, (size_t)std::max(A->nsamp, 1) * (size_t)std::max(A->gtotal, 1), h->stream);
    fill0(A->gstate, (size_t)h->ncell, h->stream);
    fill0(A->ghmax, (size_t)h->ncell, h->stream);

    int rc2 = run_captured(h, A, 0, enqueue_forward);
    if (rc2) return rc2;
    hipStream_t s = h->stream;
    std::vector<InstCtrl> c(h->ninst);
    HIPCHK(hipMemcpyAsync(c.data(), h->ctrl, c.size() * sizeof(InstCtrl), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    h->hctrl = c;
    for (int b = 0; b < h->ninst; b++) {
        h->last_run[b] = 0;
        h->last_total[b] = 0;
        if (c[b].steps != nsteps) {
            fprintf(stderr, "cuteflow_hip: recorded run of instance %d stopped after %d of %d steps\n", b,
                    c[b].steps, nsteps);
            return c[b].status < 0 ? CF_ERR_DT : CF_ERR_ARG;
        }
    }
    HIPCHK(hipMemcpyAsync(A->man_rec, h->manning, (size_t)h->ncell * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    A->params_ver = h->params_ver;
    A->recorded = true;
    A->tape_last = true;
    // the seeds were zeroed on the device above
    A->geta_host.assign((size_t)std::max(A->nsamp, 1) * (size_t)std::max(A->gtotal, 1), 0.0);
    A->geta_dirty = false;
    return CF_OK;
}

extern "C" int cf_adj_num_samples(cf_handle *h) { return (h && h->adj && h->adj->recorded) ? h->adj->nsamp : 0; }

static bool bad_adj(cf_handle *h, int inst) {
    return !h || !h->adj || !h->adj->recorded || inst < 0 || inst >= h->ninst || h->pending >= 0;
}

// Copies a per-cell internal-order device array of instance b to a caller
// pointer (host or device) in the caller's order.
static int out_scalar(cf_handle *h, int b, const double *src, double *dst) {
    const int c0 = h->cell_off[b], n = h->nelt[b];
    hipLaunchKernelGGL(scalar_out_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, src, h->perm, c0, n,
                       h->stage);
    HIPCHK(hipMemcpyAsync(dst, h->stage, sizeof(double) * (size_t)n, hipMemcpyDefault, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

static int in_scalar(cf_handle *h, int b, const double *src, double *dst) {
    const int c0 = h->cell_off[b], n = h->nelt[b];
    HIPCHK(hipMemcpyAsync(h->stage, src, sizeof(double) * (size_t)n, hipMemcpyDefault, h->stream));
    hipLaunchKernelGGL(scalar_in_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, h->stage, h->perm, c0, n,
                       dst);
    HIPCHK(hipGetLastError());
    return CF_OK;
}

static int out_state3(cf_handle *h, int b, const double4 *src, double *dst) {
    const int c0 = h->cell_off[b], n = h->nelt[b];
    hipLaunchKernelGGL(state3_out_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, src, h->perm, c0, n,
                       h->stage);
    HIPCHK(hipMemcpyAsync(dst, h->stage, sizeof(double) * 3 * (size_t)n, hipMemcpyDefault, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

static int in_state3(cf_handle *h, int b, const double *src, double4 *dst) {
    const int c0 = h->cell_off[b], n = h->nelt[b];
    HIPCHK(hipMemcpyAsync(h->stage, src, sizeof(double) * 3 * (size_t)n, hipMemcpyDefault, h->stream));
    hipLaunchKernelGGL(state3_in_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, h->stage, h->perm, c0, n,
                       dst);
    HIPCHK(hipGetLastError());
    return CF_OK;
}

// The (nsamp x ngauges) block of instance inst of a sample-major array.
// Columns of instance inst from a sample-major device array (nsamp x
// gtotal).  `cache`, when given, holds the host copy of src across calls
// (`ok`: already copied).
static int get_samples(cf_handle *h, int inst, const double *src, double *dst, std::vector<double> *cache = nullptr,
                       bool *ok = nullptr) {
    AdjState *A = h->adj;
    if (A->nsamp <= 0 || A->gtotal <= 0) return CF_OK;
    const int ng = A->goff[inst + 1] - A->goff[inst];
    std::vector<double> local;
    std::vector<double> &all = cache ? *cache : local;
    if (!cache || !*ok) {
        all.resize((size_t)A->nsamp * A->gtotal);
        HIPCHK(hipMemcpyAsync(all.data(), src, all.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (ok) *ok = true;
    }
    for (int s = 0; s < A->nsamp; s++)
        for (int g = 0; g < ng; g++) dst[(size_t)s * ng + g] = all[(size_t)s * A->gtotal + A->goff[inst] + g];
    return CF_OK;
}

extern "C" int cf_adj_get_samples(cf_handle *h, int inst, double *eta, double *hmax) {
    if (bad_adj(h, inst)) return CF_ERR_ARG;
    AdjState *A = h->adj;
    if (eta) {
        int rc = get_samples(h, inst, A->eta, eta, &A->eta_host, &A->eta_host_ok);
        if (rc) return rc;
    }
    if (hmax) {
        if (!A->want_hmax) return CF_ERR_ARG;
        return out_scalar(h, inst, A->hmax, hmax);
    }
    return CF_OK;
}

extern "C" int cf_adj_set_seed(cf_handle *h, int inst, const double *g_eta, const double *g_hmax,
                               const double *g_state) {
    if (bad_adj(h, inst)) return CF_ERR_ARG;
    AdjState *A = h->adj;
    if (g_hmax && !A->want_hmax) return CF_ERR_ARG;
    const int c0 = h->cell_off[inst], n = h->nelt[inst];
    const int ng = A->goff[inst + 1] - A->goff[inst];
    if (ng > 0 && A->nsamp > 0) {
        // seeds of this instance's gauges: strided columns of the sample-major
        // host copy, uploaded by the next reverse sweep
        std::vector<double> &all = A->geta_host;
        for (int s = 0; s < A->nsamp; s++)
            for (int g = 0; g < ng; g++)
                all[(size_t)s * A->gtotal + A->goff[inst] + g] = g_eta ? g_eta[(size_t)s * ng + g] : 0.0;
        A->geta_dirty = true;
    }
    if (A->want_hmax) {
        if (g_hmax) {
            int rc = in_scalar(h, inst, g_hmax, A->ghmax);
            if (rc) return rc;
        } else {
            fill0(A->ghmax + c0, (size_t)n, h->stream);
        }
    }
    if (g_state) {
        int rc = in_state3(h, inst, g_state, A->gstate);
        if (rc) return rc;
    } else {
        fill0(A->gstate + c0, (size_t)n, h->stream);
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

template <int IFLUX, int FRIC>
static void launch_rev_t(const RevArgs &r, int grid, hipStream_t s) {
    hipLaunchKernelGGL((revg_kernel<IFLUX, FRIC>), dim3(grid), dim3(r.tile_cells), 0, s, r, r.kp);
}

static void launch_rev(const RevArgs &r, int iflux, int fric, int grid, hipStream_t s) {
    typedef void (*fn)(const RevArgs &, int, hipStream_t);
    static const fn tab[4][4] = {{launch_rev_t<1, 0>, launch_rev_t<1, 1>, launch_rev_t<1, 2>, launch_rev_t<1, 3>},
                                 {launch_rev_t<2, 0>, launch_rev_t<2, 1>, launch_rev_t<2, 2>, launch_rev_t<2, 3>},
                                 {launch_rev_t<3, 0>, launch_rev_t<3, 1>, launch_rev_t<3, 2>, launch_rev_t<3, 3>},
                                 {launch_rev_t<4, 0>, launch_rev_t<4, 1>, launch_rev_t<4, 2>, launch_rev_t<4, 3>}};
    tab[iflux - 1][fric](r, grid, s);
}

// Reverse sweep: segments from the last, each replayed from its checkpoint
// and reversed step by step (revg_kernel: gather of step k + 1 and reverse
// of step k in one launch), then the gather of step 0 gives the cotangent
// of the initial state in lam[0].
static void enqueue_backward(cf_handle *h, AdjState *A, hipStream_t s) {
    const size_t N = (size_t)h->ncell;
    const int S = A->nsteps + 1;
    const int *am = A->want_hmax ? A->argmax : nullptr;
    // lambda of the final state: its seed + the seeds of the last sample
    const int last = (A->nsteps % A->every == 0) ? A->nsteps / A->every - 1 : -1;
    hipLaunchKernelGGL(lam_init_kernel, dim3(h->ntiles), dim3(h->tile), 0, s, h->tile_inst, h->inst_end, A->gstate, am,
                       A->ghmax, last, A->lam[1]);
    if (last >= 0 && A->gtotal)
        hipLaunchKernelGGL(gauge_seed_kernel, dim3(grid_of(h->ninst, 64)), dim3(64), 0, s, A->d_gcell, A->d_goff,
                           A->geta, A->gtotal, last, h->ninst, A->lam[1]);
    fill0(A->gn, N, s);

    StepArgs a = base_args(h);
    a.manning = A->man_rec;
    a.dt_fix = A->dt;
    a.tc_fix = A->tc;
    a.hist_stride = S;
    a.skip = 0;
    a.tile_ran = A->tile_ran;
    RevArgs r{};
    r.nbr = h->nbr;
    r.dfc = h->dfc;
    r.geo = h->geo;
    r.manning = A->man_rec;
    r.hydro = h->hydro;
    r.kp = h->d_kp;
    r.tile_inst = h->tile_inst;
    r.inst_end = h->inst_end;
    r.lam = A->lam[1];
    r.r2_off = A->r2_off;
    r.r2_idx = A->r2_idx;
    r.gn = A->gn;
    r.dt = A->dt;
    r.tc = A->tc;
    r.argmax = am;
    r.ghmax = A->ghmax;
    r.gfirst = A->d_gfirst;
    r.gnext = A->d_gnext;
    r.hstride = S;
    r.ncell = h->ncell;
    r.tile_cells = h->tile;
    r.xcd = h->xcd;
    r.tile_map = h->tile_map;
    for (int sg = A->nseg - 1; sg >= 0; sg--) {
        const int k0 = sg * A->K, k1 = std::min(k0 + A->K, A->nsteps);
        const double4 *ck = A->ckpt + (size_t)sg * N;
        // replay: state k0+m+1 -> seg[m], residual of step k0+m -> xs[m].
        // seg[m] last held step k0+K+m of the segment after this one (none
        // for the first segment reversed: its content is unknown).  After a
        // recorded forward, seg and xs hold the last segment's tape.
        const bool taped = sg == A->nseg - 1 && A->tape_last;
        if (!taped) {
            const bool has_next = sg + 1 < A->nseg;
            hipLaunchKernelGGL(replay_action_kernel, dim3(grid_of(h->ntiles, 256)), dim3(256), 0, s,
                               A->tile_ran + (size_t)k0 * h->ntiles, k1 - k0, h->ntiles,
                               has_next ? A->tile_ran + (size_t)(k0 + A->K) * h->ntiles : nullptr,
                               has_next ? std::min(k1 - k0, A->nsteps - k0 - A->K) : 0, A->act);
        }
        for (int k = k0; k < k1 && !taped; k++) {
            a.step = k;
            // workgroups to the tiles that ran at one of the next steps first,
            // in XCD-contiguous chunks whatever the forward's schedule: the
            // replay and the reverse gain from the L2 locality (the gradient of
            // a 16 x 1M-cell batch 2364 -> 2307 ms, its forward equal either way)
            if ((k - k0) % kReorderEvery == 0)
                launch_rank(h, s, A->tile_ran + (size_t)k * h->ntiles, std::min(kReorderEvery, k1 - k), 1);
            a.in_x = (k == k0) ? ck : A->seg + (size_t)(k - k0 - 1) * N;
            a.out_x = A->seg + (size_t)(k - k0) * N;
            a.xs = A->xs + (size_t)(k - k0) * 3 * N;
            a.replay_act = A->act + (size_t)(k - k0) * h->ntiles;
            launch_step(a, h->iflux, h->fric, h->ntiles, s);
        }
        for (int k = k1 - 1; k >= k0; k--) {
            r.step = k;
            if ((k1 - 1 - k) % kReorderEvery == 0) {
                const int lo = std::max(k0, k - kReorderEvery + 1);
                launch_rank(h, s, A->tile_ran + (size_t)lo * h->ntiles, k - lo + 1, 1);
            }
            r.st = (k == k0) ? ck : A->seg + (size_t)(k - k0 - 1) * N;
            r.xs = A->xs + (size_t)(k - k0) * 3 * N;
            r.ran = A->tile_ran + (size_t)k * h->ntiles;
            r.ran_next = k + 1 < A->nsteps ? A->tile_ran + (size_t)(k + 1) * h->ntiles : nullptr;
            r.R1in = A->R1[(k + 1) & 1];
            r.R2in = A->R2[(k + 1) & 1];
            r.cmask_in = A->cmask[(k + 1) & 1];
            r.cmask = A->cmask[k & 1];
            r.R1 = A->R1[k & 1];
            r.R2 = A->R2[k & 1];
            // seeds of the state after step k (sample sidx)
            r.sidx = ((k + 1) % A->every == 0) ? (k + 1) / A->every - 1 : -1;
            r.geta_row = (r.sidx >= 0 && A->gtotal) ? A->geta + (size_t)r.sidx * A->gtotal : nullptr;
            launch_rev(r, h->iflux, h->fric, h->ntiles, s);
        }
    }
    if (h->iflux >= 3)
        hipLaunchKernelGGL(gather_kernel<true>, dim3(h->ntiles), dim3(h->tile), 0, s, h->nbr, A->R1[0], A->R2[0],
                           A->r2_off, A->r2_idx, h->tile_inst, h->inst_end, A->tile_ran, A->cmask[0], A->lam[0]);
    else
        hipLaunchKernelGGL(gather_kernel<false>, dim3(h->ntiles), dim3(h->tile), 0, s, h->nbr, A->R1[0], A->R2[0],
                           A->r2_off, A->r2_idx, h->tile_inst, h->inst_end, A->tile_ran, A->cmask[0], A->lam[0]);
}

extern "C" int cf_adj_backward(cf_handle *h) {
    if (!h || !h->adj || !h->adj->recorded || h->pending >= 0) return CF_ERR_ARG;
    AdjState *A = h->adj;
    if (A->params_ver != h->params_ver) return CF_ERR_ARG;  // inflow or hydrograph changed since the record
    if (A->geta_dirty) {
        HIPCHK(hipMemcpyAsync(A->geta, A->geta_host.data(), A->geta_host.size() * sizeof(double),
                              hipMemcpyHostToDevice, h->stream));
        A->geta_dirty = false;
    }
    int rc = run_captured(h, A, A->tape_last ? 3 : 1, enqueue_backward);
    if (rc) return rc;
    // the replays of the other segments overwrote the tape
    if (A->nseg > 1) A->tape_last = false;
    HIPCHK(hipStreamSynchronize(h->stream));
    A->lam_cur = 0;  // the buffer holding the initial-state cotangent
    return CF_OK;
}

extern "C" int cf_adj_get_gradient(cf_handle *h, int inst, double *g_manning, double *g_state0) {
    if (bad_adj(h, inst)) return CF_ERR_ARG;
    AdjState *A = h->adj;
    if (g_manning) {
        int rc = out_scalar(h, inst, A->gn, g_manning);
        if (rc) return rc;
    }
    if (g_state0) return out_state3(h, inst, A->lam[A->lam_cur], g_state0);
    return CF_OK;
}

extern "C" int cf_tlm_set_direction(cf_handle *h, int inst, const double *dn, const double *dU0) {
    if (bad_adj(h, inst)) return CF_ERR_ARG;
    AdjState *A = h->adj;
    const size_t N = (size_t)h->ncell;
    if (!A->dman) {
        drop_adj_graphs(A);
        if (!dalloc(A->dman, N) || !dalloc(A->tst[0], N) || !dalloc(A->tst[1], N) || !dalloc(A->tdst[0], N) ||
            !dalloc(A->tdst[1], N) || !dalloc(A->tdir, N))
            return CF_ERR_NOMEM;
        fill0(A->dman, N, h->stream);
        fill0(A->tdir, N, h->stream);
    }
    const int c0 = h->cell_off[inst], n = h->nelt[inst];
    int rc;
    if (dn) {
        if ((rc = in_scalar(h, inst, dn, A->dman)) != CF_OK) return rc;
    } else {
        fill0(A->dman + c0, (size_t)n, h->stream);
    }
    if (dU0) {
        if ((rc = in_state3(h, inst, dU0, A->tdir)) != CF_OK) return rc;
    } else {
        fill0(A->tdir + c0, (size_t)n, h->stream);
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    A->tlm_ready = false;
    return CF_OK;
}

template <int IFLUX, int FRIC, bool REC>
static void launch_tlm_t(const TlmArgs &t, int grid, hipStream_t s) {
    hipLaunchKernelGGL((tlm_kernel<IFLUX, FRIC, REC>), dim3(grid), dim3(t.tile_cells), 0, s, t);
}

// Tangent-linear run over the recorded dt and tc from the recorded initial
// state.
static void enqueue_tlm(cf_handle *h, AdjState *A, hipStream_t s) {
    const size_t N = (size_t)h->ncell;
    typedef void (*fn)(const TlmArgs &, int, hipStream_t);
    static const fn tab[2][4][4] = {
        {{launch_tlm_t<1, 0, false>, launch_tlm_t<1, 1, false>, launch_tlm_t<1, 2, false>, launch_tlm_t<1, 3, false>},
         {launch_tlm_t<2, 0, false>, launch_tlm_t<2, 1, false>, launch_tlm_t<2, 2, false>, launch_tlm_t<2, 3, false>},
         {launch_tlm_t<3, 0, false>, launch_tlm_t<3, 1, false>, launch_tlm_t<3, 2, false>, launch_tlm_t<3, 3, false>},
         {launch_tlm_t<4, 0, false>, launch_tlm_t<4, 1, false>, launch_tlm_t<4, 2, false>, launch_tlm_t<4, 3, false>}},
        {{launch_tlm_t<1, 0, true>, launch_tlm_t<1, 1, true>, launch_tlm_t<1, 2, true>, launch_tlm_t<1, 3, true>},
         {launch_tlm_t<2, 0, true>, launch_tlm_t<2, 1, true>, launch_tlm_t<2, 2, true>, launch_tlm_t<2, 3, true>},
         {launch_tlm_t<3, 0, true>, launch_tlm_t<3, 1, true>, launch_tlm_t<3, 2, true>, launch_tlm_t<3, 3, true>},
         {launch_tlm_t<4, 0, true>, launch_tlm_t<4, 1, true>, launch_tlm_t<4, 2, true>, launch_tlm_t<4, 3, true>}}};
    // masks only for a run of the step count they were sized for
    int mode = A->mask_steps == A->nsteps ? A->mask_mode : 0;
    if (mode >= 3 && A->sig_steps != A->nsteps) mode = 0;
    const fn L = tab[mode >= 3 ? 1 : 0][h->iflux - 1][h->fric];
    // start: the recorded initial state and the direction's initial tangent
    hipLaunchKernelGGL(copy_d4_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, A->ckpt, A->tst[0], N);
    hipLaunchKernelGGL(copy_d4_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, A->tdir, A->tdst[0], N);
    TlmArgs t{};
    t.nbr = h->nbr;
    t.dfc = h->dfc;
    t.geo = h->geo;
    t.manning = h->manning;
    t.dman = A->dman;
    t.hydro = h->hydro;
    t.kp = h->d_kp;
    t.tile_inst = h->tile_inst;
    t.inst_end = h->inst_end;
    t.dt = A->dt;
    t.tc = A->tc;
    t.hstride = A->nsteps + 1;
    t.ncell = h->ncell;
    t.tile_cells = h->tile;
    t.xcd = h->xcd;
    t.tile_map = h->tile_map;
    t.mask_mode = mode;
    t.wmask = A->wmask;
    t.flips = A->flips;
    t.bsig = A->bsig;
    t.bflips = A->bflips;
    double4 *vb[2] = {A->tst[0], A->tst[1]};
    double4 *db[2] = {A->tdst[0], A->tdst[1]};
    if (A->want_hmax) fill0(A->dhmax, N, s);
    for (int k = 0; k < A->nsteps; k++) {
        t.step = k;
        t.st_in = vb[k & 1];
        t.dst_in = db[k & 1];
        t.st_out = vb[(k + 1) & 1];
        t.dst_out = db[(k + 1) & 1];
        L(t, h->ntiles, s);
        if ((k + 1) % A->every == 0) {
            const int sidx = (k + 1) / A->every - 1;
            if (A->gtotal)
                hipLaunchKernelGGL(tlm_sample_kernel, dim3(grid_of(A->gtotal, 256)), dim3(256), 0, s,
                                   vb[(k + 1) & 1], db[(k + 1) & 1], A->d_gcell, A->gtotal, sidx, A->teta, A->deta);
            if (A->want_hmax)
                hipLaunchKernelGGL(tlm_hmax_kernel, dim3(h->ntiles), dim3(h->tile), 0, s, db[(k + 1) & 1], h->tile_inst,
                                   h->inst_end, A->argmax, sidx, A->dhmax);
        }
    }
}

extern "C" int cf_tlm_set_mask_mode(cf_handle *h, int mode) {
    if (!h || !h->adj || !h->adj->recorded || h->pending >= 0 || mode < 0 || mode > 4) return CF_ERR_ARG;
    AdjState *A = h->adj;
    // counting needs the masks (and for mode 4 the signatures) of a run of
    // this step count
    if ((mode == 2 || mode == 4) && (!A->mask_rec || A->mask_steps != A->nsteps)) return CF_ERR_ARG;
    if (mode == 4 && (!A->sig_rec || A->sig_steps != A->nsteps)) return CF_ERR_ARG;
    const size_t N = (size_t)h->ncell;
    if ((mode == 1 || mode == 3) && A->mask_steps != A->nsteps) {
        dfree(A->wmask);
        A->mask_steps = 0;
        if (!dalloc(A->wmask, (size_t)A->nsteps * (N >> 6))) return CF_ERR_NOMEM;
        A->mask_steps = A->nsteps;
    }
    if (mode == 3 && A->sig_steps != A->nsteps) {
        dfree(A->bsig);
        A->sig_steps = 0;
        if (!dalloc(A->bsig, (size_t)A->nsteps * N)) return CF_ERR_NOMEM;
        A->sig_steps = A->nsteps;
    }
    if (mode && !A->flips) {
        if (!dalloc(A->flips, N) || !dalloc(A->bflips, N)) {
            dfree(A->flips);
            dfree(A->bflips);
            return CF_ERR_NOMEM;
        }
        hipLaunchKernelGGL(fill_i_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->stream, A->flips, N);
        hipLaunchKernelGGL(fill_i_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->stream, A->bflips, N);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    if (mode != A->mask_mode) drop_adj_graphs(A);
    A->mask_mode = mode;
    return CF_OK;
}

extern "C" int cf_tlm_get_flips(cf_handle *h, int inst, int kind, int *flips, int reset) {
    if (bad_adj(h, inst) || !h->adj->flips || !flips || kind < 0 || kind > 1) return CF_ERR_ARG;
    AdjState *A = h->adj;
    int *src = kind ? A->bflips : A->flips;
    const int c0 = h->cell_off[inst], n = h->nelt[inst];
    int *stage = reinterpret_cast<int *>(h->stage);  // 3 * max cells doubles
    hipLaunchKernelGGL(int_out_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, src, h->perm, c0, n, stage);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(flips, stage, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (reset)
        hipLaunchKernelGGL(fill_i_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, src + c0, (size_t)n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

extern "C" int cf_tlm_run(cf_handle *h) {
    if (!h || !h->adj || !h->adj->recorded || !h->adj->dman || h->pending >= 0) return CF_ERR_ARG;
    AdjState *A = h->adj;
    int rc = run_captured(h, A, 2, enqueue_tlm);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    A->tlm_cur = A->nsteps & 1;
    A->tlm_ready = true;
    if (A->mask_mode == 1 || A->mask_mode == 3) A->mask_rec = true;
    if (A->mask_mode == 3) A->sig_rec = true;
    return CF_OK;
}

extern "C" int cf_tlm_get(cf_handle *h, int inst, double *eta, double *d_eta, double *d_hmax, double *d_state,
                          double *state) {
    if (bad_adj(h, inst) || !h->adj->tlm_ready) return CF_ERR_ARG;
    AdjState *A = h->adj;
    int rc;
    if (eta && (rc = get_samples(h, inst, A->teta, eta)) != CF_OK) return rc;
    if (d_eta && (rc = get_samples(h, inst, A->deta, d_eta)) != CF_OK) return rc;
    if (d_hmax) {
        if (!A->want_hmax) return CF_ERR_ARG;
        if ((rc = out_scalar(h, inst, A->dhmax, d_hmax)) != CF_OK) return rc;
    }
    double4 *db[2] = {A->tdst[0], A->tdst[1]};
    double4 *vb[2] = {A->tst[0], A->tst[1]};
    if (d_state && (rc = out_state3(h, inst, db[A->tlm_cur], d_state)) != CF_OK) return rc;
    if (state && (rc = out_state3(h, inst, vb[A->tlm_cur], state)) != CF_OK) return rc;
    return CF_OK;
}
