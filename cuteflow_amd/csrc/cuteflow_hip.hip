// MI355X-native finite-volume St-Venant time step (FP64, unstructured
// triangles) behind the C ABI of include/cuteflow_hip.h.
//
// One fused kernel per time step replaces the reference's compute_flux,
// compute_source_terms, update_sol_with_flux_and_source, cflcond and
// reduction kernels (flux_riemann.cuf:4-429, source_terms.cuf:4-213,
// finite_volumes_method.cuf:260-305, cfl_condition.cuf:4-104):
//   - one thread per cell gathers its three faces (no atomics, no
//     intermediate residual arrays in HBM);
//   - the new state goes to the other half of a ping-pong buffer (no
//     vdlg0 <- vdlg copy, finite_volumes_method.cuf:230);
//   - the CFL dt of the new state is reduced with wavefront shuffles and a
//     workgroup LDS pass to one plain store per workgroup; a finalize kernel
//     (one workgroup per instance) reduces those and publishes dt, tc and
//     the stop/error status in a device-resident control block, so a run of
//     steps needs no host synchronisation and no device-wide fence or atomic
//     (on gfx950 an agent-scope fence writes back and invalidates the XCD's
//     L2);
//   - cells are stored along a Hilbert curve (256-cell tiles are compact
//     patches) and tiles that are dry with dry neighbours are skipped;
//   - a handle holds one or more instances (meshes of any size, each padded
//     to whole tiles, each with its own parameters, dt, time and status):
//     one launch per step covers all of them, and a run of steps is
//     captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "cf_device.h"
#include "cf_internal.h"
#include "cf_prep.h"

#pragma clang fp contract(off)

using namespace cf;

namespace {

__device__ __forceinline__ int nb4_at(const int4 &v, int j) { return j == 0 ? v.x : (j == 1 ? v.y : v.z); }

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = dmin(v, __shfl_xor(v, off, 64));
    return v;
}

// Occupancy target of the step kernel (waves per SIMD): 4 waves caps the
// kernel at 128 VGPRs without spilling (measured best on MI355X; 6 and 8
// spill); 0 lets the compiler choose the register budget.
// Neighbour gathers of the step kernel through LDS (TileLd) when 1.  Off:
// on MI355X staging the tile (one barrier, a range check per gather) made
// the 2M-cell reach slower (70.8 against 66.4 us/step with 128-cell tiles,
// 67.9 against 64.0 with 64-cell tiles): the neighbour reads hit L2 and are
// not what bounds the kernel.
#ifndef CF_STEP_LDS
#define CF_STEP_LDS 0
#endif

#ifndef CF_STEP_WAVES_PER_EU
#define CF_STEP_WAVES_PER_EU 4
#endif
#if CF_STEP_WAVES_PER_EU > 0
#define CF_STEP_OCCUPANCY __attribute__((amdgpu_waves_per_eu(CF_STEP_WAVES_PER_EU, 8)))
#else
#define CF_STEP_OCCUPANCY
#endif

// FRIC: 0 no friction, 1 explicit (fricimplic 0), 2 fricimplic 1, 3 fricimplic 2
// `kp` is a.kp passed again as a kernel parameter: `__restrict__` only takes
// effect on parameters, and with it the compiler keeps the per-instance
// constants in scalar registers (scalar loads) instead of reloading them per
// lane after the first store; that also frees vector registers (fewer spills).
// MODE kStepRun: a step of a run.  kStepReplay: the replay of a recorded
// run (a.dt_fix set): the recorded dt and tile activity, states from a.in_x
// to a.out_x and the residuals to a.xs.  kStepRunTape: a step of a run that
// also writes what the replay would (a.out_x, a.xs).  One kernel per mode, so
// that no mode carries another's code.
enum { kStepRun = 0, kStepReplay = 1, kStepRunTape = 2 };

template <int IFLUX, int FRIC, int MODE>
__global__ __launch_bounds__(kTile) CF_STEP_OCCUPANCY void step_kernel(StepArgs a, const KParams *__restrict__ kp) {
    __shared__ double s_min[kWaves];
    const int2 tb = a.tile_map[blockIdx.x];  // (tile, instance)
    const int tile = tb.x, b = tb.y;
    const InstCtrl c = a.ctrl[b];
    constexpr bool replay = MODE == kStepReplay, tape = MODE == kStepRunTape;
    // the tile's byte in this step's activity row of a recorded run (none
    // otherwise), addressed where it is used: off the prologue's load chain
    const auto ran_at = [&]() { return a.tile_ran + (size_t)a.step * gridDim.x + tile; };
    if (!replay) {
        // A tile whose cells and neighbour tiles' cells were all dry after
        // the last two steps would recompute (0, 0, 0) for every cell (all
        // faces dry/dry, no source, clamp), and both state buffers already
        // hold that: it is skipped.  Reading a stamp raised concurrently only
        // makes the tile run, which is always correct.  The stamp and the
        // always flag are read beside the control block and one branch takes
        // both exits (this instance stopped or failed; the tile skipped), so
        // that the loads form one round after the map rather than a chain:
        // most workgroups end here.
        const unsigned stamp = a.stamp[tile];
        const bool always = a.always[tile] != 0;
        const bool stopped = c.status != 0;
        // bitwise, not short-circuit: every load is used on every path
        if (stopped | ((a.skip != 0) & !always & (stamp + 1u < c.gstep))) {
            if (threadIdx.x == 0) {
                if (a.tile_ran) *ran_at() = 0;  // no dt contribution
                if (a.tile_class && !stopped) a.tile_class[tile] = kTileSkipped;
            }
            if (tape && !stopped) {
                const int ic = tile * (int)blockDim.x + threadIdx.x;
                a.out_x[ic] = make_double4(0.0, 0.0, 0.0, a.zm[ic]);
            }
            return;
        }
    } else {
        // skipped in the recorded run: dry cells stay (+0, +0, +0, zm); the
        // output buffer already holds that unless the tile ran at the step
        // it last held (or that step is unknown): one byte says which
        const int act = a.replay_act[tile];
        if (act != kReplayRun) {
            if (act == kReplayRewrite) {
                const int ic = tile * (int)blockDim.x + threadIdx.x;
                a.out_x[ic] = make_double4(0.0, 0.0, 0.0, a.zm[ic]);
            }
            return;
        }
    }
    const double4 *__restrict__ st_in = replay ? a.in_x : a.st[(c.par + a.step) & 1];
    double4 *__restrict__ st_out = replay ? a.out_x : a.st[(c.par + a.step + 1) & 1];
    const int i = tile * (int)blockDim.x + threadIdx.x;
#if CF_STEP_LDS
    // the tile's states and neighbour codes in LDS (blockDim.x cells each)
    extern __shared__ double4 s_tile[];
    double4 *s_st = s_tile;
    int4 *s_nb = reinterpret_cast<int4 *>(s_tile + blockDim.x);
    s_st[threadIdx.x] = st_in[i];
    s_nb[threadIdx.x] = a.nbr[i];
    __syncthreads();
#endif
    const size_t hix = (size_t)b * a.hist_stride + a.step;
    const double dt = replay ? a.dt_fix[hix] : c.dt;
    const double tc = replay ? a.tc_fix[hix] : c.tc;
    const KParams &p = kp[b];
    const double tls = p.tolisec;
    double dti = 1000000.0;
    int isnan_local = 0, wet_local = 0;

    if (i < a.inst_end[b]) {
#if CF_STEP_LDS
        const TileLd ld{st_in, a.nbr, a.dfc, a.geo, a.manning, a.ncell, s_st, s_nb, tile * (int)blockDim.x,
                        (int)blockDim.x};
        const double zml = s_st[threadIdx.x].w;
#else
        const DevLd ld{st_in, a.nbr, a.dfc, a.geo, a.manning, a.ncell};
        const double zml = st_in[i].w;
#endif
        double o[3], xo[3];
        cell_update<double, IFLUX, FRIC>(ld, p, a.hydro, i, dt, tc, o, MODE != kStepRun ? xo : nullptr,
                                         a.surf_in ? a.surf_in + (size_t)b * CF_MAX_IO : nullptr);
        const double h = o[0], hu = o[1], hv = o[2];
        st_out[i] = make_double4(h, hu, hv, zml);
        if (replay || tape) {
            if (tape) a.out_x[i] = make_double4(h, hu, hv, zml);
            if (a.xs) {
                a.xs[i] = xo[0];
                a.xs[(size_t)a.ncell + i] = xo[1];
                a.xs[2 * (size_t)a.ncell + i] = xo[2];
            }
            if (replay) return;
        }
        isnan_local = (h != h) || (hu != hu) || (hv != hv);
        wet_local = (h != 0.0) || isnan_local;
        // cflcond on the new state (cfl_condition.cuf:20-36)
        if (h > tls) {
            const double uu = hu / h, vv = hv / h;
            const double nv = sqrt(uu * uu + vv * vv);
            const double vmax = dmax(0.01, nv + sqrt(p.gp * fabs(h)));
            dti = dmin((p.cfl * a.geo[G_COT * (size_t)a.ncell + i]) / vmax, 1000000.0);
        }
        if (p.is_dt_constant == 1) dti = p.constant_dt;
    }
    if (replay) return;
    // per-workgroup dt minimum: wavefront shuffles, then LDS across waves;
    // then one device-scope atomic min into one of the instance's slots (the
    // minima are positive doubles, whose bit patterns order like their
    // values, so the unsigned minimum of the bits is the exact minimum) and
    // one count; finalize_kernel reads the slots after the step
    const double m = wave_min(dti);
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = m;
    const int anynan = __syncthreads_or(isnan_local);
    if (threadIdx.x == 0) {
        double bm = s_min[0];
#pragma unroll
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) bm = dmin(bm, s_min[w]);
        const int slot = (int)(blockIdx.x & (kDtSlots - 1));
        const size_t so = (size_t)b * kDtSlots + slot;
        if (anynan) atomicOr(a.nanflag + (size_t)b * kSlotStride32, 1u);
        else atomicMin(a.dtslot + so * kSlotStride64, (unsigned long long)__double_as_longlong(bm));
        atomicAdd(a.cntslot + so * kSlotStride32, 1u);
        // stored last: an early store would keep the uniform loads above
        // from being scalar loads (they could alias it)
        if (a.tile_ran) *ran_at() = 1;
    }
    // a wet tile keeps itself and its neighbour tiles running for two steps
    // and is ranked among the first tiles of the next steps
    const int wet = (a.skip || a.tile_class) ? __syncthreads_or(wet_local) : 0;
    if (a.skip && wet) {
        const int o = a.tile_adj_off[tile], n = a.tile_adj_off[tile + 1] - o;
        for (int k = threadIdx.x; k <= n; k += (int)blockDim.x)
            a.stamp[k == 0 ? tile : a.tile_adj[o + k - 1]] = c.gstep + 1u;
    }
    if (a.tile_class && threadIdx.x == 0) a.tile_class[tile] = wet ? kTileWet : kTileDry;
}

// The step kernel's workgroup -> (tile, instance) map of the wet-first
// order, in two launches of 256-tile blocks.  A stable partition of the
// tiles by class (wet, dry that ran, skipped: tile_class, written by the
// step kernel) gives each class a run of workgroups; with xcd set the tiles
// of a class are cut into one contiguous chunk per XCD (workgroup b runs on
// XCD b mod 8, which keeps a chunk's neighbour reads in that XCD's L2),
// sized by the number of the class's workgroups on each XCD.  Every tile
// gets exactly one workgroup.  Class counts travel packed, 21 bits each
// (ntiles < 2^21).
constexpr int kRankBlock = 256;

// A tile's class: its tile_class entry, or with `nrows` activity rows of a
// recorded run (ran[j * ntiles + t] = 1: tile t ran at that step), wet for
// the tiles that ran at one of those steps and skipped for the others.
__device__ __forceinline__ int class_of(const unsigned char *__restrict__ tile_class,
                                        const unsigned char *__restrict__ ran, int nrows, int t, int ntiles) {
    if (!ran) return min((int)tile_class[t], 2);
    // independent loads (no early exit: a chain of dependent loads is slower)
    int any = 0;
#pragma unroll 8
    for (int j = 0; j < nrows; j++) any |= ran[(size_t)j * ntiles + t];
    return any ? kTileWet : kTileSkipped;
}

__device__ __forceinline__ unsigned long long class_bit(const unsigned char *__restrict__ tile_class,
                                                       const unsigned char *__restrict__ ran, int nrows, int t,
                                                       int ntiles) {
    return t < ntiles ? 1ull << (21 * class_of(tile_class, ran, nrows, t, ntiles)) : 0ull;
}

// inclusive scan of packed counts over the workgroup (kRankBlock threads)
__device__ __forceinline__ unsigned long long block_scan_packed(unsigned long long v, unsigned long long &total) {
    __shared__ unsigned long long s_w[kRankBlock / 64];
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    if (lane == 63) s_w[w] = v;
    __syncthreads();
    unsigned long long before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < kRankBlock / 64; k++) {
        before += k < w ? s_w[k] : 0ull;
        all += s_w[k];
    }
    total = all;
    return v + before;
}

// counts[k]: the packed class counts of tiles 256k .. 256k + 255
__global__ __launch_bounds__(kRankBlock) void rank_count_kernel(const unsigned char *__restrict__ tile_class,
                                                                const unsigned char *__restrict__ ran, int nrows,
                                                                int ntiles, unsigned long long *__restrict__ counts) {
    unsigned long long total;
    (void)block_scan_packed(class_bit(tile_class, ran, nrows, (int)(blockIdx.x * kRankBlock + threadIdx.x), ntiles),
                            total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

__global__ __launch_bounds__(kRankBlock) void rank_map_kernel(const unsigned char *__restrict__ tile_class,
                                                              const unsigned char *__restrict__ ran, int nrows,
                                                              const int *__restrict__ tile_inst, int ntiles, int xcd,
                                                              const unsigned long long *__restrict__ counts,
                                                              int2 *__restrict__ map) {
    __shared__ unsigned long long s_sum[2][kRankBlock / 64];
    const int nb = (int)gridDim.x, k0 = (int)blockIdx.x, lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    // the counts of the blocks before this one, and of all blocks
    unsigned long long before = 0, all = 0;
    for (int k = (int)threadIdx.x; k < nb; k += kRankBlock) {
        const unsigned long long c = counts[k];
        all += c;
        if (k < k0) before += c;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        before += __shfl_xor(before, off, 64);
        all += __shfl_xor(all, off, 64);
    }
    if (lane == 0) {
        s_sum[0][w] = before;
        s_sum[1][w] = all;
    }
    __syncthreads();
    before = 0;
    all = 0;
#pragma unroll
    for (int q = 0; q < kRankBlock / 64; q++) {
        before += s_sum[0][q];
        all += s_sum[1][q];
    }
    const int t = k0 * kRankBlock + (int)threadIdx.x;
    const unsigned long long mine = class_bit(tile_class, ran, nrows, t, ntiles);
    unsigned long long total;
    const unsigned long long incl = block_scan_packed(mine, total);
    if (t >= ntiles) return;
    constexpr unsigned long long kM = (1ull << 21) - 1;
    const int c = class_of(tile_class, ran, nrows, t, ntiles);
    const int j = (int)(((before + incl - mine) >> (21 * c)) & kM);  // rank of tile t within its class
    const int n0 = (int)(all & kM), n1 = (int)((all >> 21) & kM), n2 = (int)((all >> 42) & kM);
    const int p = c == 0 ? 0 : (c == 1 ? n0 : n0 + n1), n = c == 0 ? n0 : (c == 1 ? n1 : n2);
    int blk = p + j;
    if (xcd) {
        // the class's workgroups p .. p+n-1 on XCD x: n/8, one more for the
        // first n%8 XCDs in cyclic order from p's; chunk x of the class's
        // tiles goes to them in increasing order
        int off = 0;
        for (int x = 0; x < 8; x++) {
            const int cx = n / 8 + ((((x - p) & 7) < (n & 7)) ? 1 : 0);
            if (j < off + cx) {
                blk = p + ((x - p) & 7) + 8 * (j - off);
                break;
            }
            off += cx;
        }
    }
    map[blk] = make_int2(t, tile_inst[t]);
}

// The fixed workgroup -> (tile, instance) map: the identity, or with xcd set
// one contiguous run of tiles per XCD (tile_of_block's order).
__global__ void static_map_kernel(const int *__restrict__ tile_inst, int ntiles, int xcd, int2 *__restrict__ map) {
    const int bid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (bid >= ntiles) return;
    int t = bid;
    if (xcd) {
        const int x = bid & 7, q = ntiles >> 3, r = ntiles & 7;
        t = x * q + (x < r ? x : r) + (bid >> 3);
    }
    map[bid] = make_int2(t, tile_inst[t]);
}

// Reduces per-tile minima (NaN marks a NaN state, +inf a skipped tile)
// starting from `pad`; every thread gets the result, the NaN flag and the
// number of skipped tiles.  Each thread issues its loads eight at a time
// before using them: one workgroup reads up to ~16k partials, and a loop
// that waits for each load in turn is bound by memory latency.
__device__ double block_min_1024(const double *__restrict__ part, int n, double pad, int &nan, int &nskip) {
    __shared__ double s[1024 / 64];
    __shared__ int sk[1024 / 64];
    constexpr int kU = 8;
    const int stride = (int)blockDim.x;
    double m = pad;
    int bad = 0, skipped = 0;
    for (int base = threadIdx.x; base < n; base += kU * stride) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int k = base + u * stride;
            v[u] = k < n ? part[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            if (base + u * stride < n) {
                if (v[u] != v[u]) bad = 1;
                else m = dmin(m, v[u]);
                skipped += (v[u] == __builtin_inf());
            }
        }
    }
    m = wave_min(m);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) skipped += __shfl_xor(skipped, off, 64);
    if ((threadIdx.x & 63) == 0) {
        s[threadIdx.x >> 6] = m;
        sk[threadIdx.x >> 6] = skipped;
    }
    nan = __syncthreads_or(bad);
    double r = s[0];
    nskip = sk[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
        r = dmin(r, s[w]);
        nskip += sk[w];
    }
    return r;
}

// Discharges of a step, one thread per inflow/outflow cell (slot s, cell
// io_cell[s]): cell_discharge() on the step's input state, which the step
// kernel leaves intact (it writes the other state buffer).  Runs between
// the step kernel and finalize_kernel, only with discharge bookkeeping on.
__global__ void discharge_kernel(StepArgs a, const KParams *__restrict__ kp, const int *__restrict__ io_cell, int nio,
                                 double *__restrict__ q_io) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= nio) return;
    const int i = io_cell[sl];
    const int b = a.tile_inst[i / a.tile_cells];
    const InstCtrl c = a.ctrl[b];
    if (c.status != 0) return;
    const DevLd ld{a.st[(c.par + a.step) & 1], a.nbr, a.dfc, a.geo, a.manning, a.ncell};
    double qo, qi;
    cell_discharge(ld, kp[b], a.hydro, i, c.tc, qo, qi, a.surf_in ? a.surf_in + (size_t)b * CF_MAX_IO : nullptr);
    q_io[2 * (size_t)sl] = qo;
    q_io[2 * (size_t)sl + 1] = qi;
}

// phys_inflow (calcul_surf_entree, maillage.cuf:1368-1421, run before each
// step by finite_volumes_method.cuf:133-139): thread id-1 of workgroup b
// sums, over instance b's cells with an inflow or outflow face in the
// caller's cell order, the length of the cell's last face on inlet id times
// the cell's depth in the step's input state (the reference assigns per
// face, so a cell's last face on the inlet counts; it adds block sums with
// atomics, in no fixed order).
__global__ void surf_in_kernel(StepArgs a, const int *__restrict__ io_cell, const int *__restrict__ io_off,
                               double *__restrict__ surf_in) {
    const int b = blockIdx.x, id = (int)threadIdx.x + 1;
    if (id > CF_MAX_IO) return;
    const InstCtrl c = a.ctrl[b];
    const double4 *__restrict__ st = a.st[(c.par + a.step) & 1];
    double s = 0.0;
    for (int q = io_off[b]; q < io_off[b + 1]; q++) {
        const int i = io_cell[q];
        const int4 nb = a.nbr[i];
        double v = 0.0;
        bool hit = false;
        for (int j = 0; j < 3; j++) {
            const int code = nb4_at(nb, j);
            if (code < 0 && ((-code) & 3) == 1 && ((-code) >> 2) == id) {
                v = a.geo[(G_LN + j) * (size_t)a.ncell + i] * st[i].x;
                hit = true;
            }
        }
        if (hit) s = s + v;
    }
    surf_in[(size_t)b * CF_MAX_IO + id - 1] = s;
}

// End of a step, one workgroup per instance: dt of the new state (the
// reference's reduction, cfl_condition.cuf:68-104: the step kernel's tiles
// have folded their minima into the instance's slots, which are read here
// and emptied for the next step), tc += dt, the dt range check (:129-132)
// and the loop condition (finite_volumes_method.cuf:123).  With discharge
// bookkeeping (q_io set): the step's inflow and outflow discharges (summed
// over the instance's slots in the caller's cell order) and
// check_steady_state (finite_volumes_method.cuf:370-393): more than 1000
// consecutive steps with |Qin - Qout| / Qin < tol_reg_perm stop the
// instance after the step.  mirror (may be null): the run record's copy of
// the control blocks, kept equal to ctrl.
__global__ __launch_bounds__(256) void finalize_kernel(InstCtrl *ctrl, const KParams *kp,
                                                        unsigned long long *dtslot, unsigned *cntslot,
                                                        unsigned *nanflag, double *dt_hist, double *tc_hist, int *ran,
                                                        int hist_stride, int step, const double *q_io,
                                                        const int *io_off, double *qin_hist, double *qout_hist,
                                                        InstCtrl *mirror) {
    __shared__ double s_q[2 * 256];
    const int b = blockIdx.x;
    // every load up front (one memory round trip): the control block and the
    // constants (uniform), the slots (lanes 0..kDtSlots-1)
    InstCtrl cc = ctrl[b];
    const double pad = kp[b].pad_dt, ts = kp[b].ts, tol = kp[b].tol, trp = kp[b].tol_reg_perm;
    const unsigned nanv = nanflag[(size_t)b * kSlotStride32];
    unsigned long long bits = kDtSlotEmpty;
    unsigned cnt = 0;
    const size_t so = (size_t)b * kDtSlots + threadIdx.x;
    if (threadIdx.x < kDtSlots) {
        bits = dtslot[so * kSlotStride64];
        cnt = cntslot[so * kSlotStride32];
    }
    // a stopped instance's tiles did not run: its slots are still empty
    if (cc.status != 0) return;
    if (threadIdx.x < kDtSlots) {
        dtslot[so * kSlotStride64] = kDtSlotEmpty;
        cntslot[so * kSlotStride32] = 0u;
    }
    double qo = 0.0, qi = 0.0;
    if (q_io) {
        const int s0 = io_off[b], s1 = io_off[b + 1];
        for (int base = s0; base < s1; base += (int)blockDim.x) {
            const int n = min((int)blockDim.x, s1 - base);
            __syncthreads();
            if ((int)threadIdx.x < n) {
                s_q[2 * threadIdx.x] = q_io[2 * (size_t)(base + threadIdx.x)];
                s_q[2 * threadIdx.x + 1] = q_io[2 * (size_t)(base + threadIdx.x) + 1];
            }
            __syncthreads();
            if (threadIdx.x == 0)
                for (int k = 0; k < n; k++) {
                    qo = qo + s_q[2 * k];
                    qi = qi + s_q[2 * k + 1];
                }
        }
    }
    if (threadIdx.x < 64) {
#pragma unroll
        for (int off = kDtSlots / 2; off >= 1; off >>= 1) {
            const unsigned long long o = __shfl_xor(bits, off, 64);
            bits = o < bits ? o : bits;
            cnt += __shfl_xor(cnt, off, 64);
        }
    }
    if (threadIdx.x == 0) {
        const bool nan = nanv != 0u;
        if (nan) nanflag[(size_t)b * kSlotStride32] = 0u;
        const double dtn = dmin(pad, __longlong_as_double((long long)bits));
        const double tcn = cc.tc + cc.dt;
        cc.dt = dtn;
        cc.tc = tcn;
        cc.steps = step + 1;
        cc.gstep += 1u;
        dt_hist[(size_t)b * hist_stride + step + 1] = dtn;
        tc_hist[(size_t)b * hist_stride + step + 1] = tcn;
        ran[(size_t)b * hist_stride + step] = (int)cnt;
        bool steady = false;
        if (q_io) {
            qout_hist[(size_t)b * hist_stride + step] = qo;
            qin_hist[(size_t)b * hist_stride + step] = qi;
            if (qi > kQSteady && fabs(qi - qo) / qi < trp) steady = ++cc.reg > 1000;
            else cc.reg = 0;
        }
        if (nan || !(dtn <= 9999.0 && dtn >= kDtLow)) cc.status = -1;
        else if (steady) cc.status = 2;
        else if (!(((tcn - ts) + 2.0 * tol) < tol)) cc.status = 1;
        ctrl[b] = cc;
        if (mirror) mirror[b] = cc;
    }
}

// Start of a run: the buffer parity moves past the previous run's steps,
// the loop condition is evaluated (a failed instance stays failed, one at
// steady state stays stopped: a new state restarts both).
__global__ void run_init_kernel(InstCtrl *ctrl, const KParams *kp, double *dt_hist, double *tc_hist,
                                int hist_stride, int ninst, InstCtrl *mirror) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ninst) return;
    InstCtrl *c = ctrl + b;
    c->par = (c->par + c->steps) & 1;
    c->steps = 0;
    if (c->status == 0 || c->status == 1) c->status = (((c->tc - kp[b].ts) + 2.0 * kp[b].tol) < kp[b].tol) ? 0 : 1;
    dt_hist[(size_t)b * hist_stride] = c->dt;
    tc_hist[(size_t)b * hist_stride] = c->tc;
    if (mirror) mirror[b] = *c;
}

// cflcond for the current state of one instance, one minimum per tile
// (cfl_condition.cuf:4-39); like the reference, the initial dt does not
// look for NaN states.
__global__ __launch_bounds__(kTile) void cfl_kernel(const double4 *__restrict__ st, const double *__restrict__ cot,
                                                     int cell0, int cell_end, const KParams *kp, int b,
                                                     double *__restrict__ part) {
    __shared__ double s_min[kWaves];
    const int i = cell0 + blockIdx.x * (int)blockDim.x + threadIdx.x;
    const KParams &p = kp[b];
    double dti = 1000000.0;
    if (i < cell_end) {
        const double4 u = st[i];
        const double h = u.x;
        if (h > p.tolisec) {
            const double uu = u.y / h, vv = u.z / h;
            const double nv = sqrt(uu * uu + vv * vv);
            const double vmax = dmax(0.01, nv + sqrt(p.gp * fabs(h)));
            dti = dmin((p.cfl * cot[i]) / vmax, 1000000.0);
        }
        if (p.is_dt_constant == 1) dti = p.constant_dt;
    }
    const double m = wave_min(dti);
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double bm = s_min[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) bm = dmin(bm, s_min[w]);
        part[blockIdx.x] = bm;
    }
}

// dt of the current state (finite_volumes_method.cuf:113-116); a new state
// restarts the instance's clock.
__global__ __launch_bounds__(1024) void reset_inst_kernel(const double *part, int nparts, const KParams *kp, int b,
                                                           InstCtrl *ctrl) {
    int nan, nskip;
    const double m = block_min_1024(part, nparts, kp[b].pad_dt, nan, nskip);
    if (threadIdx.x == 0) {
        ctrl[b].dt = nan ? __builtin_nan("") : m;
        ctrl[b].tc = 0.0;
        ctrl[b].status = 0;
        ctrl[b].reg = 0;
    }
}

// Geometry of each cell (maillage.cuf:132-159, :1254-1331, flux_riemann.cuf:50-74).
__global__ void geometry_kernel(const int *__restrict__ conn, const double *__restrict__ coords, int n,
                                int cell0, int stride, double *__restrict__ geo, double *__restrict__ zm_out) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int i = cell0 + l;
    const int s1 = conn[3 * l] - 1, s2 = conn[3 * l + 1] - 1, s3 = conn[3 * l + 2] - 1;
    const double vx[3] = {coords[3 * s1], coords[3 * s2], coords[3 * s3]};
    const double vy[3] = {coords[3 * s1 + 1], coords[3 * s2 + 1], coords[3 * s3 + 1]};
    const double vz[3] = {coords[3 * s1 + 2], coords[3 * s2 + 2], coords[3 * s3 + 2]};
    double g[G_COUNT], zm;
    cell_geometry(vx, vy, vz, g, zm);
    const size_t S = (size_t)stride;
#pragma unroll
    for (int k = 0; k < G_COUNT; k++) geo[k * S + i] = g[k];
    zm_out[i] = zm;
}

// Centroid distance across each face (0 across a boundary).
__global__ void face_dist_kernel(const int4 *__restrict__ nbr, const double *__restrict__ geo, int cell0,
                                 int cell_end, int stride, double4 *__restrict__ dfc) {
    const int i = cell0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cell_end) return;
    const size_t S = (size_t)stride;
    const int4 nb = nbr[i];
    const double xi = geo[G_XC * S + i], yi = geo[G_YC * S + i];
    double d[3];
    for (int j = 0; j < 3; j++) {
        const int k = nb4_at(nb, j);
        const int kk = k >= 0 ? k : i;
        const double xk = geo[G_XC * S + kk], yk = geo[G_YC * S + kk];
        d[j] = sqrt((xi - xk) * (xi - xk) + (yi - yk) * (yi - yk));
    }
    dfc[i] = make_double4(d[0], d[1], d[2], 0.0);
}

// State in the caller's order and layout (3, nelt) <-> internal double4
// (h, hu, hv, zm) in the internal cell order; perm: internal -> caller.
__global__ void pack_state_kernel(const double *__restrict__ U, const double *__restrict__ zm,
                                  const int *__restrict__ perm, int cell0, int n, double4 *__restrict__ st) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) {
        const int i = cell0 + l;
        const size_t o = 3 * (size_t)perm[i];
        st[i] = make_double4(U[o], U[o + 1], U[o + 2], zm[i]);
    }
}

__global__ void unpack_state_kernel(const double4 *__restrict__ st, const int *__restrict__ perm, int cell0, int n,
                                    double *__restrict__ U) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) {
        const int i = cell0 + l;
        const double4 s = st[i];
        const size_t o = 3 * (size_t)perm[i];
        U[o] = s.x;
        U[o + 1] = s.y;
        U[o + 2] = s.z;
    }
}

// per-cell scalar in the caller's order -> internal cell order
__global__ void gather_scalar_kernel(const double *__restrict__ src, const int *__restrict__ perm, int cell0, int n,
                                     double *__restrict__ dst) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) dst[cell0 + l] = src[perm[cell0 + l]];
}

__global__ void fill_u32_kernel(unsigned *p, int n, unsigned v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void pow_kernel(const double *__restrict__ x, double y, int n, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cf_pow(x[i], y);
}

}  // namespace

static KParams make_kparams(const cf_params &p, int nelt, int hydro_off) {
    KParams k{};
    k.gp = p.gp; k.tolisec = p.tolisec; k.tol = p.tol; k.cfl = p.cfl; k.ts = p.ts;
    k.constant_dt = p.constant_dt; k.override_manning = p.override_manning;
    k.pad_dt = (nelt % 1024 != 0) ? 100000.0 : 1000000.0;
    k.tol_reg_perm = p.tol_reg_perm;
    k.ilimiteur = p.ilimiteur; k.inlet = p.inlet; k.outlet = p.outlet;
    k.is_dt_constant = p.is_dt_constant; k.is_override_manning = p.is_override_manning;
    k.hydro_off = hydro_off;
    for (int i = 0; i < CF_MAX_IO; i++) {
        k.debitglob[i] = p.debitglob[i];
        k.h_sortie[i] = p.h_sortie[i];
        k.long_entree[i] = p.long_entree[i];
        k.hydro_n[i] = p.hydro_n[i];
    }
    return k;
}

static constexpr int kHydroPerInst = 2 * CF_MAX_IO * CF_MAX_HYDRO;

static int upload_params(cf_handle *h, int b) {
    std::vector<double> buf(kHydroPerInst);
    memcpy(buf.data(), h->prm[b].hydro_t, sizeof(double) * CF_MAX_IO * CF_MAX_HYDRO);
    memcpy(buf.data() + CF_MAX_IO * CF_MAX_HYDRO, h->prm[b].hydro_q, sizeof(double) * CF_MAX_IO * CF_MAX_HYDRO);
    HIPCHK(hipMemcpyAsync(h->hydro + (size_t)b * kHydroPerInst, buf.data(), buf.size() * sizeof(double),
                          hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_kp + b, &h->kp[b], sizeof(KParams), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

static void drop_graphs(cf_handle *h) {
    for (auto &g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
}

// dt of instance b's current state; restarts its clock.
static int compute_dt(cf_handle *h, int b) {
    const int c0 = h->cell_off[b], nt = h->tile_off[b + 1] - h->tile_off[b];
    const int cur = (h->hctrl[b].par + h->hctrl[b].steps) & 1;
    double *part = h->partial + h->tile_off[b];
    hipLaunchKernelGGL(cfl_kernel, dim3(nt), dim3(h->tile), 0, h->stream, h->st[cur], h->geo + G_COT * (size_t)h->ncell,
                       c0, c0 + h->nelt[b], h->d_kp, b, part);
    hipLaunchKernelGGL(reset_inst_kernel, dim3(1), dim3(1024), 0, h->stream, part, nt, h->d_kp, b, h->ctrl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->hctrl.data() + b, h->ctrl + b, sizeof(InstCtrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

static void free_record(cf_handle *h) {
    dfree(h->rec);
    if (h->hrec) (void)hipHostFree(h->hrec);
    h->hrec = nullptr;
    h->ctrl_mirror = nullptr;
    h->dt_hist = h->tc_hist = h->qin_hist = h->qout_hist = nullptr;
    h->ran = nullptr;
}

// Sizes the run record for runs of up to nsteps steps (history rows of
// cap + 1 entries per instance).
static int ensure_hist(cf_handle *h, int nsteps) {
    if (nsteps <= h->cap) return CF_OK;
    const int cap = std::max(nsteps, 2 * h->cap);
    HIPCHK(hipStreamSynchronize(h->stream));
    free_record(h);
    drop_graphs(h);
    const size_t n = (size_t)h->ninst * (size_t)(cap + 1), B = (size_t)h->ninst;
    h->rec_dt = round256(B * sizeof(InstCtrl));
    h->rec_ran = h->rec_dt + round256(n * sizeof(double));
    h->rec_qin = h->rec_ran + round256(n * sizeof(int));
    h->rec_qout = h->rec_qin + round256(n * sizeof(double));
    h->rec_tc = h->rec_qout + round256(n * sizeof(double));
    h->rec_bytes = h->rec_tc + round256(n * sizeof(double));
    if (!dalloc(h->rec, h->rec_bytes) || hipHostMalloc((void **)&h->hrec, h->rec_tc, hipHostMallocDefault) != hipSuccess) {
        h->hrec = nullptr;
        free_record(h);
        return CF_ERR_NOMEM;
    }
    h->ctrl_mirror = reinterpret_cast<InstCtrl *>(h->rec);
    h->dt_hist = reinterpret_cast<double *>(h->rec + h->rec_dt);
    h->ran = reinterpret_cast<int *>(h->rec + h->rec_ran);
    h->qin_hist = reinterpret_cast<double *>(h->rec + h->rec_qin);
    h->qout_hist = reinterpret_cast<double *>(h->rec + h->rec_qout);
    h->tc_hist = reinterpret_cast<double *>(h->rec + h->rec_tc);
    h->cap = cap;
    return CF_OK;
}

__global__ void fill_u64_kernel(unsigned long long *p, size_t n, unsigned long long v) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

extern "C" const char *cf_error_string(int s) {
    switch (s) {
    case CF_OK: return "ok";
    case CF_ERR_ARG: return "invalid argument or unsupported option";
    case CF_ERR_HIP: return "HIP runtime error";
    case CF_ERR_DT: return "time step out of range (dt > 9999, dt < 1e-12) or NaN state";
    case CF_ERR_MESH: return "inconsistent mesh";
    case CF_ERR_NOMEM: return "out of memory";
    case CF_ERR_IO: return "file missing, unreadable or not in the expected layout";
    case CF_ERR_SIZE: return "input larger than the fixed scratch of a no-allocation entry point";
    default: return "unknown status";
    }
}

static int validate_params(const cf_params *p) {
    if (!(p->gp > 0.0) || !(p->cfl > 0.0) || !(p->tolisec >= 0.0)) return CF_ERR_ARG;
    if (p->iflux < 1 || p->iflux > 4) return CF_ERR_ARG;
    if (p->inlet < 1 || p->inlet > 3) return CF_ERR_ARG;
    if (p->outlet != 1 && p->outlet != 2) return CF_ERR_ARG;
    if (p->friction != 0 && p->friction != 1) return CF_ERR_ARG;
    if (p->fricimplic < 0 || p->fricimplic > 2) return CF_ERR_ARG;
    if (p->is_dt_constant == 1 && !(p->constant_dt > 0.0)) return CF_ERR_ARG;
    if (p->nombre_entrees < 0 || p->nombre_entrees > CF_MAX_IO) return CF_ERR_ARG;
    for (int i = 0; i < CF_MAX_IO; i++)
        if (p->hydro_n[i] < 0 || p->hydro_n[i] > CF_MAX_HYDRO) return CF_ERR_ARG;
    return CF_OK;
}

static void destroy_buffers(cf_handle *h) {
    drop_graphs(h);
    dfree(h->st[0]); dfree(h->st[1]); dfree(h->nbr); dfree(h->geo); dfree(h->dfc);
    dfree(h->manning); dfree(h->hydro); dfree(h->zm); dfree(h->stage); dfree(h->perm);
    dfree(h->d_kp); dfree(h->ctrl); dfree(h->partial); dfree(h->stamp);
    dfree(h->tile_adj_off); dfree(h->tile_adj); dfree(h->always);
    dfree(h->tile_inst); dfree(h->inst_end); dfree(h->inst_tile);
    dfree(h->tile_map); dfree(h->tile_class); dfree(h->rank_counts);
    free_record(h);
    dfree(h->dtslot); dfree(h->cntslot); dfree(h->nanflag);
    dfree(h->io_cell); dfree(h->io_off); dfree(h->q_io);
    dfree(h->surf_in);
}

extern "C" int cf_batch_create(int ninst, const cf_mesh *meshes, const cf_params *params, cf_handle **out) {
    if (ninst <= 0 || !meshes || !params || !out) return CF_ERR_ARG;
    *out = nullptr;
    for (int b = 0; b < ninst; b++) {
        const cf_mesh *m = meshes + b;
        if (m->nelt <= 0 || m->nnt <= 0 || !m->conn || !m->coords || !m->boundary) return CF_ERR_ARG;
        if ((size_t)m->nelt > (size_t)INT32_MAX / 8) return CF_ERR_ARG;
        int st = validate_params(params + b);
        if (st) return st;
        // one kernel serves the whole batch: same scheme and friction mode
        if (params[b].iflux != params[0].iflux || params[b].friction != params[0].friction ||
            params[b].fricimplic != params[0].fricimplic || params[b].options != params[0].options)
            return CF_ERR_ARG;
        for (size_t i = 0; i < 3 * (size_t)m->nelt; i++)
            if (m->conn[i] < 1 || m->conn[i] > m->nnt) return CF_ERR_MESH;
    }
    if ((params[0].options & ~(CF_OPT_KEEP_ORDER | CF_OPT_NO_TILE_SKIP | CF_OPT_TILE_128 | CF_OPT_TILE_64 |
                               CF_OPT_XCD_CONTIG | CF_OPT_FIXED_TILE_ORDER)) ||
        ((params[0].options & CF_OPT_TILE_128) && (params[0].options & CF_OPT_TILE_64)))
        return CF_ERR_ARG;
    const int tile = (params[0].options & CF_OPT_TILE_64) ? 64 : (params[0].options & CF_OPT_TILE_128) ? 128 : kTile;
    // Instances that pass the same mesh arrays (e.g. scenarios of one reach)
    // share one preparation: it depends on the mesh alone, apart from the
    // inflow-length check of inlet 2.
    std::vector<cfprep::Prep> prep(ninst);
    for (int b = 0; b < ninst; b++) {
        const cf_mesh &m = meshes[b];
        int same = -1;
        for (int q = 0; q < b && same < 0; q++) {
            const cf_mesh &mq = meshes[q];
            if (mq.nelt == m.nelt && mq.nnt == m.nnt && mq.conn == m.conn && mq.coords == m.coords &&
                mq.boundary == m.boundary && params[q].inlet == params[b].inlet &&
                memcmp(params[q].long_entree, params[b].long_entree, sizeof(params[b].long_entree)) == 0)
                same = q;
        }
        if (same >= 0) {
            prep[b] = prep[same];
            continue;
        }
        int st = cfprep::prepare_device(meshes + b, params + b, tile, (params[b].options & CF_OPT_KEEP_ORDER) != 0,
                                        prep[b]);
        if (st) return st;
    }
    cf_handle *h = new (std::nothrow) cf_handle();
    if (!h) return CF_ERR_NOMEM;
    h->ninst = ninst;
    h->tile = tile;
    h->iflux = params[0].iflux;
    h->fric = params[0].friction == 0 ? 0 : 1 + params[0].fricimplic;
    h->nelt.resize(ninst); h->nnt.resize(ninst); h->cell_off.resize(ninst); h->tile_off.resize(ninst + 1);
    h->prm.assign(params, params + ninst);
    h->kp.resize(ninst);
    h->hperm.resize(ninst);
    h->hctrl.assign(ninst, InstCtrl{});
    h->last_run.assign(ninst, 0);
    h->last_total.assign(ninst, 0);
    h->last_qin.assign(ninst, {});
    h->last_qout.assign(ninst, {});
    // skipping dry tiles is exact except with a constant dt (every cell then
    // contributes constant_dt to the minimum)
    h->skip = (params[0].options & CF_OPT_NO_TILE_SKIP) ? 0 : 1;
    for (int b = 0; b < ninst; b++)
        if (params[b].is_dt_constant == 1) h->skip = 0;
    h->xcd = (params[0].options & CF_OPT_XCD_CONTIG) ? 1 : 0;
    h->wet_first = (params[0].options & CF_OPT_FIXED_TILE_ORDER) ? 0 : 1;
    for (int b = 0; b < ninst; b++)
        if (params[b].inlet == 3) h->phys_in = 1;
    int cells = 0, tiles = 0;
    for (int b = 0; b < ninst; b++) {
        h->nelt[b] = meshes[b].nelt;
        h->nnt[b] = meshes[b].nnt;
        h->cell_off[b] = cells;
        h->tile_off[b] = tiles;
        const int nt = grid_of(meshes[b].nelt, tile);
        tiles += nt;
        cells += nt * tile;
        h->max_cells = std::max(h->max_cells, meshes[b].nelt);
        h->kp[b] = make_kparams(params[b], meshes[b].nelt, b * kHydroPerInst);
        h->hperm[b] = prep[b].perm;
    }
    h->tile_off[ninst] = tiles;
    h->ncell = cells;
    h->ntiles = tiles;
    if (tiles >= (1 << 21)) h->wet_first = 0;  // reorder_kernel packs tile counts in 21 bits
    const size_t N = (size_t)cells;

    // global (batched) host tables
    std::vector<int4> nbr(N, make_int4(-3, -3, -3, 0));
    std::vector<double> mann(N, 0.0);
    std::vector<int> perm(N, 0), tile_inst(tiles), inst_end(ninst), inst_tile(ninst + 1), adj_off(tiles + 1, 0), adj;
    std::vector<unsigned> always(tiles, 0u);
    // discharge slots: the instances' inflow/outflow cells, caller order
    std::vector<int> io_cell, io_off(ninst + 1, 0);
    for (int b = 0; b < ninst; b++) {
        io_off[b + 1] = io_off[b] + (int)prep[b].io_cells.size();
        for (size_t q = 0; q < prep[b].io_cells.size(); q++) io_cell.push_back(h->cell_off[b] + prep[b].io_cells[q]);
        if (params[b].tol_reg_perm > 0.0) h->discharge = 1;
    }
    h->nio = io_off[ninst];
    for (int b = 0; b < ninst; b++) {
        const int c0 = h->cell_off[b], t0 = h->tile_off[b];
        for (int i = 0; i < h->nelt[b]; i++) {
            const cfprep::Nbr &r = prep[b].nbr[i];
            int c[3];
            for (int k = 0; k < 3; k++) c[k] = r.c[k] >= 0 ? r.c[k] + c0 : r.c[k];
            nbr[c0 + i] = make_int4(c[0], c[1], c[2], r.w);
            perm[c0 + i] = prep[b].perm[i];
            if (meshes[b].manning) mann[c0 + i] = meshes[b].manning[prep[b].perm[i]];
        }
        inst_end[b] = c0 + h->nelt[b];
        inst_tile[b] = t0;
        const int nt = h->tile_off[b + 1] - t0;
        for (int t = 0; t < nt; t++) {
            tile_inst[t0 + t] = b;
            always[t0 + t] = prep[b].tile_always[t];
            for (int q = prep[b].tile_off[t]; q < prep[b].tile_off[t + 1]; q++) adj.push_back(prep[b].tile_adj[q] + t0);
            adj_off[t0 + t + 1] = (int)adj.size();
        }
    }
    inst_tile[ninst] = tiles;

    int rc = CF_OK;
    std::vector<int *> dconn(ninst, nullptr);
    std::vector<double *> dcoords(ninst, nullptr);
    do {
        if (hipStreamCreate(&h->stream) != hipSuccess) { rc = CF_ERR_HIP; break; }
        h->own_stream = true;
        if (!dalloc(h->st[0], N) || !dalloc(h->st[1], N) || !dalloc(h->nbr, N) || !dalloc(h->geo, G_COUNT * N) ||
            !dalloc(h->dfc, N) || !dalloc(h->manning, N) || !dalloc(h->hydro, (size_t)ninst * kHydroPerInst) ||
            !dalloc(h->zm, N) || !dalloc(h->stage, 3 * (size_t)h->max_cells) || !dalloc(h->perm, N) ||
            !dalloc(h->d_kp, (size_t)ninst) || !dalloc(h->ctrl, (size_t)ninst) || !dalloc(h->partial, (size_t)tiles) ||
            !dalloc(h->stamp, (size_t)tiles) || !dalloc(h->tile_adj_off, adj_off.size()) ||
            !dalloc(h->tile_adj, adj.size()) || !dalloc(h->always, (size_t)tiles) ||
            !dalloc(h->tile_inst, (size_t)tiles) || !dalloc(h->inst_end, (size_t)ninst) ||
            !dalloc(h->tile_map, (size_t)tiles) || !dalloc(h->tile_class, (size_t)tiles) ||
            !dalloc(h->rank_counts, (size_t)grid_of(tiles, kRankBlock)) ||
            !dalloc(h->inst_tile, (size_t)ninst + 1) || !dalloc(h->io_cell, io_cell.size()) || !dalloc(h->io_off, (size_t)ninst + 1) ||
            !dalloc(h->q_io, 2 * std::max<size_t>((size_t)h->nio, 1)) ||
            !dalloc(h->surf_in, (size_t)ninst * CF_MAX_IO) ||
            !dalloc(h->dtslot, (size_t)ninst * kDtSlots * kSlotStride64) ||
            !dalloc(h->cntslot, (size_t)ninst * kDtSlots * kSlotStride32) ||
            !dalloc(h->nanflag, (size_t)ninst * kSlotStride32)) {
            rc = CF_ERR_NOMEM;
            break;
        }
        for (int b = 0; b < ninst && rc == CF_OK; b++)
            if (!dalloc(dconn[b], 3 * (size_t)h->nelt[b]) || !dalloc(dcoords[b], 3 * (size_t)h->nnt[b])) rc = CF_ERR_NOMEM;
        if (rc) break;
        bool ok = h2d(h->nbr, nbr.data(), N, h->stream) && h2d(h->manning, mann.data(), N, h->stream) &&
                  h2d(h->perm, perm.data(), N, h->stream) && h2d(h->tile_adj_off, adj_off.data(), adj_off.size(), h->stream) &&
                  h2d(h->tile_adj, adj.data(), adj.size(), h->stream) && h2d(h->always, always.data(), (size_t)tiles, h->stream) &&
                  h2d(h->tile_inst, tile_inst.data(), (size_t)tiles, h->stream) &&
                  h2d(h->inst_end, inst_end.data(), (size_t)ninst, h->stream) &&
                  h2d(h->inst_tile, inst_tile.data(), (size_t)ninst + 1, h->stream) &&
                  h2d(h->io_cell, io_cell.data(), io_cell.size(), h->stream) &&
                  h2d(h->io_off, io_off.data(), (size_t)ninst + 1, h->stream) &&
                  hipMemsetAsync(h->q_io, 0, 2 * std::max<size_t>((size_t)h->nio, 1) * sizeof(double), h->stream) ==
                      hipSuccess &&
                  h2d(h->ctrl, h->hctrl.data(), (size_t)ninst, h->stream) &&
                  hipMemsetAsync(h->stamp, 0, (size_t)tiles * sizeof(unsigned), h->stream) == hipSuccess &&
                  hipMemsetAsync(h->stage, 0, 3 * (size_t)h->max_cells * sizeof(double), h->stream) == hipSuccess &&
                  hipMemsetAsync(h->geo, 0, G_COUNT * N * sizeof(double), h->stream) == hipSuccess &&
                  hipMemsetAsync(h->zm, 0, N * sizeof(double), h->stream) == hipSuccess &&
                  hipMemsetAsync(h->dfc, 0, N * sizeof(double4), h->stream) == hipSuccess;
        for (int b = 0; b < ninst && ok; b++)
            ok = h2d(dconn[b], prep[b].conn.data(), 3 * (size_t)h->nelt[b], h->stream) &&
                 h2d(dcoords[b], meshes[b].coords, 3 * (size_t)h->nnt[b], h->stream);
        if (!ok) { rc = CF_ERR_HIP; break; }
        {
            const size_t n64 = (size_t)ninst * kDtSlots * kSlotStride64, n32 = (size_t)ninst * kDtSlots * kSlotStride32;
            hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_of((int)n64, 256)), dim3(256), 0, h->stream, h->dtslot, n64,
                               kDtSlotEmpty);
            // every tile starts in the first class: the first ranking keeps
            // the fixed order
            hipLaunchKernelGGL(static_map_kernel, dim3(grid_of(tiles, 256)), dim3(256), 0, h->stream, h->tile_inst,
                               tiles, h->xcd, h->tile_map);
            if (hipMemsetAsync(h->tile_class, 0, (size_t)tiles, h->stream) != hipSuccess) {
                rc = CF_ERR_HIP;
                break;
            }
            hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_of((int)n32, 256)), dim3(256), 0, h->stream, h->cntslot,
                               (int)n32, 0u);
            hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_of(ninst * kSlotStride32, 256)), dim3(256), 0, h->stream,
                               h->nanflag, ninst * kSlotStride32, 0u);
        }
        for (int b = 0; b < ninst; b++) {
            const int c0 = h->cell_off[b], n = h->nelt[b];
            hipLaunchKernelGGL(geometry_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, dconn[b], dcoords[b], n,
                               c0, h->ncell, h->geo, h->zm);
        }
        for (int b = 0; b < ninst; b++) {
            const int c0 = h->cell_off[b], n = h->nelt[b];
            hipLaunchKernelGGL(face_dist_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, h->nbr, h->geo, c0,
                               c0 + n, h->ncell, h->dfc);
            // zero state with the bed level in the 4th component, in both buffers
            for (int k = 0; k < 2; k++)
                hipLaunchKernelGGL(pack_state_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, h->stage, h->zm,
                                   h->perm, c0, n, h->st[k]);
        }
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) { rc = CF_ERR_HIP; break; }
        for (int b = 0; b < ninst && rc == CF_OK; b++) rc = upload_params(h, b);
        if (rc) break;
        if ((rc = ensure_hist(h, 1024)) != CF_OK) break;
        for (int b = 0; b < ninst && rc == CF_OK; b++) rc = compute_dt(h, b);
    } while (0);
    for (int b = 0; b < ninst; b++) {
        dfree(dconn[b]);
        dfree(dcoords[b]);
    }
    if (rc != CF_OK) {
        destroy_buffers(h);
        if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return rc;
    }
    {
        const char *env = getenv("CUTEFLOW_NO_GRAPHS");
        h->use_graphs = env && atoi(env) ? 0 : 1;
    }
    h->hnbr.swap(nbr);
    *out = h;
    return CF_OK;
}

extern "C" int cf_create(const cf_mesh *mesh, const cf_params *params, cf_handle **out) {
    return cf_batch_create(1, mesh, params, out);
}

extern "C" void cf_destroy(cf_handle *h) {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    adj_free(h);
    destroy_buffers(h);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" int cf_batch_size(cf_handle *h) { return h ? h->ninst : 0; }

static bool bad_inst(cf_handle *h, int b) { return !h || b < 0 || b >= h->ninst || h->pending >= 0; }

static int current_buffer(cf_handle *h, int b) { return (h->hctrl[b].par + h->hctrl[b].steps) & 1; }

extern "C" int cf_batch_set_state_device(cf_handle *h, int b, const double *dU) {
    if (bad_inst(h, b) || !dU) return CF_ERR_ARG;
    const int c0 = h->cell_off[b], n = h->nelt[b];
    hipLaunchKernelGGL(pack_state_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, dU, h->zm, h->perm, c0, n,
                       h->st[current_buffer(h, b)]);
    // every tile of the instance runs the next two steps (both state
    // buffers become consistent)
    const int t0 = h->tile_off[b], nt = h->tile_off[b + 1] - t0;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_of(nt, 256)), dim3(256), 0, h->stream, h->stamp + t0, nt,
                       h->hctrl[b].gstep);
    HIPCHK(hipGetLastError());
    return compute_dt(h, b);
}

extern "C" int cf_batch_set_state(cf_handle *h, int b, const double *U) {
    if (bad_inst(h, b) || !U) return CF_ERR_ARG;
    HIPCHK(hipMemcpyAsync(h->stage, U, 3 * (size_t)h->nelt[b] * sizeof(double), hipMemcpyHostToDevice, h->stream));
    return cf_batch_set_state_device(h, b, h->stage);
}

extern "C" int cf_batch_get_state_device(cf_handle *h, int b, double *dU) {
    if (bad_inst(h, b) || !dU) return CF_ERR_ARG;
    const int c0 = h->cell_off[b], n = h->nelt[b];
    hipLaunchKernelGGL(unpack_state_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream, h->st[current_buffer(h, b)],
                       h->perm, c0, n, dU);
    HIPCHK(hipGetLastError());
    return CF_OK;
}

extern "C" int cf_batch_get_state(cf_handle *h, int b, double *U) {
    if (bad_inst(h, b) || !U) return CF_ERR_ARG;
    int rc = cf_batch_get_state_device(h, b, h->stage);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(U, h->stage, 3 * (size_t)h->nelt[b] * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

extern "C" int cf_set_state_device(cf_handle *h, const double *dU) { return cf_batch_set_state_device(h, 0, dU); }
extern "C" int cf_set_state(cf_handle *h, const double *U) { return cf_batch_set_state(h, 0, U); }
extern "C" int cf_get_state_device(cf_handle *h, double *dU) { return cf_batch_get_state_device(h, 0, dU); }
extern "C" int cf_get_state(cf_handle *h, double *U) { return cf_batch_get_state(h, 0, U); }

extern "C" int cf_batch_get_time(cf_handle *h, int b, double *tc, double *dt) {
    if (!h || b < 0 || b >= h->ninst) return CF_ERR_ARG;
    if (tc) *tc = h->hctrl[b].tc;
    if (dt) *dt = h->hctrl[b].dt;
    return CF_OK;
}

extern "C" int cf_initial_dt(cf_handle *h, double *dt) {
    if (!h || !dt) return CF_ERR_ARG;
    *dt = h->hctrl[0].dt;
    return CF_OK;
}

extern "C" int cf_get_time(cf_handle *h, double *tc, double *dt) { return cf_batch_get_time(h, 0, tc, dt); }

template <int IFLUX, int FRIC>
static void launch_step_t(const StepArgs &a, int grid, hipStream_t s) {
    const size_t lds = CF_STEP_LDS ? (size_t)a.tile_cells * (sizeof(double4) + sizeof(int4)) : 0;
    if (a.dt_fix)
        hipLaunchKernelGGL((step_kernel<IFLUX, FRIC, kStepReplay>), dim3(grid), dim3(a.tile_cells), lds, s, a, a.kp);
    else if (a.out_x)
        hipLaunchKernelGGL((step_kernel<IFLUX, FRIC, kStepRunTape>), dim3(grid), dim3(a.tile_cells), lds, s, a, a.kp);
    else
        hipLaunchKernelGGL((step_kernel<IFLUX, FRIC, kStepRun>), dim3(grid), dim3(a.tile_cells), lds, s, a, a.kp);
}

typedef void (*step_launcher)(const StepArgs &, int, hipStream_t);

static step_launcher pick_launcher(int iflux, int fric) {
    static const step_launcher tab[4][4] = {
        {launch_step_t<1, 0>, launch_step_t<1, 1>, launch_step_t<1, 2>, launch_step_t<1, 3>},
        {launch_step_t<2, 0>, launch_step_t<2, 1>, launch_step_t<2, 2>, launch_step_t<2, 3>},
        {launch_step_t<3, 0>, launch_step_t<3, 1>, launch_step_t<3, 2>, launch_step_t<3, 3>},
        {launch_step_t<4, 0>, launch_step_t<4, 1>, launch_step_t<4, 2>, launch_step_t<4, 3>}};
    return tab[iflux - 1][fric];
}

void cf::launch_step(const StepArgs &a, int iflux, int fric, int grid, hipStream_t s) {
    pick_launcher(iflux, fric)(a, grid, s);
}

void cf::launch_finalize(cf_handle *h, double *dt_hist, double *tc_hist, int *ran, int stride, int step,
                         hipStream_t s, double *qin_hist, double *qout_hist, InstCtrl *mirror) {
    const bool q = qin_hist && qout_hist;
    // one wave suffices without discharges (a few loads and one thread's
    // bookkeeping); the discharge sums stage slots through LDS
    hipLaunchKernelGGL(finalize_kernel, dim3(h->ninst), dim3(q ? 256 : 64), 0, s, h->ctrl, h->d_kp, h->dtslot,
                       h->cntslot, h->nanflag, dt_hist, tc_hist, ran, stride, step, q ? h->q_io : nullptr, h->io_off,
                       qin_hist, qout_hist, mirror);
}

void cf::launch_rank(cf_handle *h, hipStream_t s, const unsigned char *ran, int nrows, int xcd) {
    if (!h->wet_first) return;
    const int nb = grid_of(h->ntiles, kRankBlock);
    hipLaunchKernelGGL(rank_count_kernel, dim3(nb), dim3(kRankBlock), 0, s, h->tile_class, ran, nrows, h->ntiles,
                       h->rank_counts);
    hipLaunchKernelGGL(rank_map_kernel, dim3(nb), dim3(kRankBlock), 0, s, h->tile_class, ran, nrows, h->tile_inst,
                       h->ntiles, xcd < 0 ? h->xcd : xcd, h->rank_counts, h->tile_map);
}

void cf::launch_run_init(cf_handle *h, double *dt_hist, double *tc_hist, int stride, hipStream_t s,
                         InstCtrl *mirror) {
    hipLaunchKernelGGL(run_init_kernel, dim3(grid_of(h->ninst, 256)), dim3(256), 0, s, h->ctrl, h->d_kp, dt_hist,
                       tc_hist, stride, h->ninst, mirror);
}

StepArgs cf::base_args(cf_handle *h) {
    StepArgs a{};
    a.st[0] = h->st[0];
    a.st[1] = h->st[1];
    a.zm = h->zm;
    a.nbr = h->nbr;
    a.geo = h->geo;
    a.dfc = h->dfc;
    a.manning = h->manning;
    a.hydro = h->hydro;
    a.kp = h->d_kp;
    a.ctrl = h->ctrl;
    a.dt_hist = h->dt_hist;
    a.tc_hist = h->tc_hist;
    a.ran = h->ran;
    a.dtslot = h->dtslot;
    a.cntslot = h->cntslot;
    a.nanflag = h->nanflag;
    a.stamp = h->stamp;
    a.tile_adj_off = h->tile_adj_off;
    a.tile_adj = h->tile_adj;
    a.always = h->always;
    a.tile_inst = h->tile_inst;
    a.inst_end = h->inst_end;
    a.inst_tile = h->inst_tile;
    a.dt_fix = nullptr;
    a.tc_fix = nullptr;
    a.in_x = nullptr;
    a.out_x = nullptr;
    a.xs = nullptr;
    a.tile_ran = nullptr;
    a.surf_in = h->phys_in ? h->surf_in : nullptr;
    a.ncell = h->ncell;
    a.tile_cells = h->tile;
    a.hist_stride = h->cap + 1;
    a.skip = h->skip;
    a.xcd = h->xcd;
    a.tile_map = h->tile_map;
    a.tile_class = h->wet_first ? h->tile_class : nullptr;
    return a;
}

// The launches of a run of nsteps steps, on stream s.
static void enqueue_run(cf_handle *h, int nsteps, hipStream_t s) {
    launch_run_init(h, h->dt_hist, h->tc_hist, h->cap + 1, s, h->ctrl_mirror);
    StepArgs a = base_args(h);
    step_launcher L = pick_launcher(h->iflux, h->fric);
    for (int k = 0; k < nsteps; k++) {
        a.step = k;
        if (k % kReorderEvery == 0) launch_rank(h, s);
        if (h->phys_in)
            hipLaunchKernelGGL(surf_in_kernel, dim3(h->ninst), dim3(32), 0, s, a, h->io_cell, h->io_off, h->surf_in);
        L(a, h->ntiles, s);
        if (h->discharge && h->nio)
            hipLaunchKernelGGL(discharge_kernel, dim3(grid_of(h->nio, 256)), dim3(256), 0, s, a, h->d_kp, h->io_cell,
                               h->nio, h->q_io);
        launch_finalize(h, h->dt_hist, h->tc_hist, h->ran, h->cap + 1, k, s, h->discharge ? h->qin_hist : nullptr,
                        h->discharge ? h->qout_hist : nullptr, h->ctrl_mirror);
    }
}

// A run's launches are captured once per step count and replayed: every
// argument is fixed (per-run state lives in the device control block).
// Finds or captures the graph of a run of nsteps steps (nothing is launched).
static int graph_of(cf_handle *h, int nsteps, hipGraphExec_t *out) {
    for (auto &g : h->graphs)
        if (g.nsteps == nsteps) {
            *out = g.exec;
            return CF_OK;
        }
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_run(h, nsteps, h->stream);
    HIPCHK(hipStreamEndCapture(h->stream, &graph));
    hipGraphExec_t exec;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    if (h->graphs.size() >= 8) {
        (void)hipGraphExecDestroy(h->graphs.front().exec);
        h->graphs.erase(h->graphs.begin());
    }
    h->graphs.push_back({nsteps, exec});
    *out = exec;
    return CF_OK;
}

// the null stream cannot be captured; very long runs are enqueued directly
static bool graphed(const cf_handle *h, int nsteps) { return h->use_graphs && h->stream && nsteps <= 4096; }

extern "C" int cf_step_prepare(cf_handle *h, int nsteps) {
    if (!h || nsteps < 0) return CF_ERR_ARG;
    if (h->pending >= 0) return CF_ERR_ARG;
    int rc = ensure_hist(h, std::max(nsteps, 1));
    if (rc) return rc;
    if (nsteps == 0 || !graphed(h, nsteps)) return CF_OK;
    hipGraphExec_t exec;
    return graph_of(h, nsteps, &exec);
}

extern "C" int cf_step_async(cf_handle *h, int nsteps) {
    if (!h || nsteps < 0) return CF_ERR_ARG;
    if (h->pending >= 0) return CF_ERR_ARG;  // previous run not finished
    bool any = false;
    for (int b = 0; b < h->ninst; b++) any = any || h->hctrl[b].status >= 0;
    if (!any) return CF_ERR_DT;
    int rc = ensure_hist(h, std::max(nsteps, 1));
    if (rc) return rc;
    if (nsteps > 0 && graphed(h, nsteps)) {
        hipGraphExec_t exec;
        rc = graph_of(h, nsteps, &exec);
        if (rc) return rc;
        h->pending = nsteps;
        HIPCHK(hipGraphLaunch(exec, h->stream));
        return CF_OK;
    }
    h->pending = nsteps;
    if (nsteps == 0) return CF_OK;
    enqueue_run(h, nsteps, h->stream);
    HIPCHK(hipGetLastError());
    return CF_OK;
}

extern "C" int cf_batch_step_finish(cf_handle *h, double *dt_hist, int *steps_done, int *status) {
    if (!h || h->pending < 0) return CF_ERR_ARG;
    const int nsteps = h->pending;
    h->pending = -1;
    const int B = h->ninst, S = h->cap + 1;
    if (nsteps == 0) {
        for (int b = 0; b < B; b++) {
            if (dt_hist) dt_hist[(size_t)b * (nsteps + 1)] = h->hctrl[b].dt;
            if (steps_done) steps_done[b] = 0;
            if (status) status[b] = h->hctrl[b].status < 0 ? CF_ERR_DT : CF_OK;
        }
        return CF_OK;
    }
    // one copy of the run record: control blocks, dt and tile counts (and
    // the discharge histories when they are kept)
    HIPCHK(hipMemcpyAsync(h->hrec, h->rec, h->discharge ? h->rec_tc : h->rec_qin, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const InstCtrl *cm = reinterpret_cast<const InstCtrl *>(h->hrec);
    std::vector<InstCtrl> c(cm, cm + B);
    const double *dts = reinterpret_cast<const double *>(h->hrec + h->rec_dt);
    const int *ran = reinterpret_cast<const int *>(h->hrec + h->rec_ran);
    const double *qin = reinterpret_cast<const double *>(h->hrec + h->rec_qin);
    const double *qout = reinterpret_cast<const double *>(h->hrec + h->rec_qout);
    int rc = CF_OK;
    for (int b = 0; b < B; b++) {
        const int done = c[b].steps;
        if (dt_hist) memcpy(dt_hist + (size_t)b * (nsteps + 1), dts + (size_t)b * S, sizeof(double) * ((size_t)done + 1));
        if (steps_done) steps_done[b] = done;
        if (status) status[b] = c[b].status < 0 ? CF_ERR_DT : CF_OK;
        if (c[b].status < 0) rc = CF_ERR_DT;
        long long run = 0;
        for (int k = 0; k < done; k++) run += ran[(size_t)b * S + k];
        h->last_run[b] = run;
        h->last_total[b] = (long long)done * (h->tile_off[b + 1] - h->tile_off[b]);
        if (h->discharge) {
            h->last_qin[b].assign(qin + (size_t)b * S, qin + (size_t)b * S + done);
            h->last_qout[b].assign(qout + (size_t)b * S, qout + (size_t)b * S + done);
        } else {
            h->last_qin[b].clear();
            h->last_qout[b].clear();
        }
    }
    h->hctrl = c;
    return rc;
}

extern "C" int cf_step_finish(cf_handle *h, double *dt_hist, int *steps_done) {
    if (!h || h->pending < 0) return CF_ERR_ARG;
    const int nsteps = h->pending;
    std::vector<double> hist((size_t)h->ninst * (nsteps + 1));
    std::vector<int> done(h->ninst), st(h->ninst);
    int rc = cf_batch_step_finish(h, hist.data(), done.data(), st.data());
    if (dt_hist) memcpy(dt_hist, hist.data(), sizeof(double) * ((size_t)done[0] + 1));
    if (steps_done) *steps_done = done[0];
    if (rc == CF_ERR_HIP) return rc;
    return st[0];
}

extern "C" int cf_step(cf_handle *h, int nsteps, double *dt_hist, int *steps_done) {
    if (steps_done) *steps_done = 0;
    int rc = cf_step_async(h, nsteps);
    if (rc) return rc;
    return cf_step_finish(h, dt_hist, steps_done);
}

extern "C" int cf_batch_step(cf_handle *h, int nsteps, double *dt_hist, int *steps_done, int *status) {
    int rc = cf_step_async(h, nsteps);
    if (rc) return rc;
    return cf_batch_step_finish(h, dt_hist, steps_done, status);
}

// Switches dry-tile skipping and the workgroup -> tile order of a handle
// (the autotuner's runtime knobs).  Captured graphs hold the old kernel
// arguments and are dropped.  Stamps are not kept while skipping is off,
// so turning it on restamps every tile to run the next two steps.
int cf::set_schedule(cf_handle *h, int skip, int xcd) {
    if (!h || h->pending >= 0) return CF_ERR_ARG;
    skip = skip ? 1 : 0;
    xcd = xcd ? 1 : 0;
    for (int b = 0; b < h->ninst; b++)
        if (h->prm[b].is_dt_constant == 1) skip = 0;
    if (skip == h->skip && xcd == h->xcd) return CF_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (skip && !h->skip)
        for (int b = 0; b < h->ninst; b++) {
            const int t0 = h->tile_off[b], nt = h->tile_off[b + 1] - t0;
            hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_of(nt, 256)), dim3(256), 0, h->stream, h->stamp + t0, nt,
                               h->hctrl[b].gstep);
        }
    if (xcd != h->xcd)
        hipLaunchKernelGGL(static_map_kernel, dim3(grid_of(h->ntiles, 256)), dim3(256), 0, h->stream, h->tile_inst,
                           h->ntiles, xcd, h->tile_map);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    drop_graphs(h);
    adj_drop_graphs(h);
    h->skip = skip;
    h->xcd = xcd;
    return CF_OK;
}

extern "C" int cf_last_run_stats(cf_handle *h, long long *tiles_run, long long *tiles_total, int *tile_cells) {
    if (!h) return CF_ERR_ARG;
    long long r = 0, t = 0;
    for (int b = 0; b < h->ninst; b++) {
        r += h->last_run[b];
        t += h->last_total[b];
    }
    if (tiles_run) *tiles_run = r;
    if (tiles_total) *tiles_total = t;
    if (tile_cells) *tile_cells = h->tile;
    return CF_OK;
}

extern "C" int cf_set_discharge(cf_handle *h, int on) {
    if (!h || h->pending >= 0) return CF_ERR_ARG;
    on = on ? 1 : 0;
    for (int b = 0; b < h->ninst; b++)
        if (h->prm[b].tol_reg_perm > 0.0) on = 1;  // the steady-state check needs it
    if (on != h->discharge) {
        HIPCHK(hipStreamSynchronize(h->stream));
        drop_graphs(h);
        h->discharge = on;
    }
    return CF_OK;
}

extern "C" int cf_batch_get_discharge(cf_handle *h, int b, double *q_in, double *q_out, int *steady) {
    if (bad_inst(h, b)) return CF_ERR_ARG;
    if (q_in) memcpy(q_in, h->last_qin[b].data(), h->last_qin[b].size() * sizeof(double));
    if (q_out) memcpy(q_out, h->last_qout[b].data(), h->last_qout[b].size() * sizeof(double));
    if (steady) *steady = h->hctrl[b].status == 2 ? 1 : 0;
    return CF_OK;
}

extern "C" int cf_get_discharge(cf_handle *h, double *q_in, double *q_out, int *steady) {
    return cf_batch_get_discharge(h, 0, q_in, q_out, steady);
}

extern "C" int cf_batch_set_inflow(cf_handle *h, int b, const double *debitglob, int n) {
    if (bad_inst(h, b) || !debitglob || n < 0 || n > CF_MAX_IO) return CF_ERR_ARG;
    for (int i = 0; i < n; i++) h->prm[b].debitglob[i] = debitglob[i];
    h->params_ver++;
    h->kp[b] = make_kparams(h->prm[b], h->nelt[b], b * kHydroPerInst);
    h->kp[b].ad_flags = h->ad_flags;
    return upload_params(h, b);
}

extern "C" int cf_set_inflow(cf_handle *h, const double *debitglob, int n) { return cf_batch_set_inflow(h, 0, debitglob, n); }

extern "C" int cf_batch_set_hydrograph(cf_handle *h, int b, int inlet, int npts, const double *t, const double *q) {
    if (bad_inst(h, b) || inlet < 1 || inlet > CF_MAX_IO || npts < 0 || npts > CF_MAX_HYDRO || (npts > 0 && (!t || !q)))
        return CF_ERR_ARG;
    for (int i = 0; i + 1 < npts; i++)
        if (!(t[i + 1] > t[i])) return CF_ERR_ARG;
    h->params_ver++;
    h->prm[b].hydro_n[inlet - 1] = npts;
    for (int i = 0; i < npts; i++) {
        h->prm[b].hydro_t[(inlet - 1) * CF_MAX_HYDRO + i] = t[i];
        h->prm[b].hydro_q[(inlet - 1) * CF_MAX_HYDRO + i] = q[i];
    }
    h->kp[b] = make_kparams(h->prm[b], h->nelt[b], b * kHydroPerInst);
    h->kp[b].ad_flags = h->ad_flags;
    return upload_params(h, b);
}

extern "C" int cf_set_hydrograph(cf_handle *h, int inlet, int npts, const double *t, const double *q) {
    return cf_batch_set_hydrograph(h, 0, inlet, npts, t, q);
}

extern "C" int cf_batch_set_manning(cf_handle *h, int b, const double *n) {
    if (bad_inst(h, b) || !n) return CF_ERR_ARG;
    const int N = h->nelt[b];
    // host or device pointer: staged, then gathered into the internal order
    HIPCHK(hipMemcpyAsync(h->stage, n, (size_t)N * sizeof(double), hipMemcpyDefault, h->stream));
    hipLaunchKernelGGL(gather_scalar_kernel, dim3(grid_of(N, 256)), dim3(256), 0, h->stream, h->stage, h->perm,
                       h->cell_off[b], N, h->manning);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

extern "C" int cf_set_manning(cf_handle *h, const double *n) { return cf_batch_set_manning(h, 0, n); }

extern "C" int cf_batch_get_geometry(cf_handle *h, int b, double *surf, double *zm, double *cotemin_arr, double *cotemin) {
    if (bad_inst(h, b)) return CF_ERR_ARG;
    const size_t N = (size_t)h->nelt[b], c0 = (size_t)h->cell_off[b], S = (size_t)h->ncell;
    std::vector<double> s(N), z(N), cot(N);
    HIPCHK(hipMemcpyAsync(s.data(), h->geo + G_SURF * S + c0, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(z.data(), h->zm + c0, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(cot.data(), h->geo + G_COT * S + c0, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double m = 1000000.0;
    for (size_t i = 0; i < N; i++) {
        const size_t o = (size_t)h->hperm[b][i];
        if (surf) surf[o] = s[i];
        if (zm) zm[o] = z[i];
        if (cotemin_arr) cotemin_arr[o] = cot[i];
        m = std::min(m, cot[i]);
    }
    if (cotemin) *cotemin = m;
    return CF_OK;
}

extern "C" int cf_get_geometry(cf_handle *h, double *surf, double *zm, double *cotemin_arr, double *cotemin) {
    return cf_batch_get_geometry(h, 0, surf, zm, cotemin_arr, cotemin);
}

extern "C" void *cf_get_stream(cf_handle *h) { return h ? (void *)h->stream : nullptr; }

extern "C" int cf_set_stream(cf_handle *h, void *stream) {
    if (!h || h->pending >= 0) return CF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
    return CF_OK;
}

extern "C" int cf_debug_prepare(const cf_mesh *mesh, const cf_params *params, int tile, int keep_order, int device,
                                int *perm, int *nbr, int *conn, int *tile_off, int *tile_adj, int adj_cap, int *n_adj,
                                unsigned char *tile_always, int *io_cells, int *n_io) {
    if (!mesh || !params || (tile != 64 && tile != 128 && tile != 256) || !perm || !nbr || !conn || !tile_off ||
        !tile_adj || !n_adj || !tile_always || !io_cells || !n_io || mesh->nelt <= 0 || !mesh->conn ||
        !mesh->coords || !mesh->boundary)
        return CF_ERR_ARG;
    cfprep::Prep P;
    const int st = device ? cfprep::prepare_device(mesh, params, tile, keep_order != 0, P)
                          : cfprep::prepare(mesh, params, tile, keep_order != 0, P);
    if (st) return st;
    if ((int)P.tile_adj.size() > adj_cap) return CF_ERR_ARG;
    const size_t N = (size_t)mesh->nelt;
    std::copy(P.perm.begin(), P.perm.end(), perm);
    for (size_t i = 0; i < N; i++) {
        for (int k = 0; k < 3; k++) nbr[4 * i + k] = P.nbr[i].c[k];
        nbr[4 * i + 3] = P.nbr[i].w;
    }
    std::copy(P.conn.begin(), P.conn.end(), conn);
    std::copy(P.tile_off.begin(), P.tile_off.end(), tile_off);
    std::copy(P.tile_adj.begin(), P.tile_adj.end(), tile_adj);
    *n_adj = (int)P.tile_adj.size();
    std::copy(P.tile_always.begin(), P.tile_always.end(), tile_always);
    std::copy(P.io_cells.begin(), P.io_cells.end(), io_cells);
    *n_io = (int)P.io_cells.size();
    return CF_OK;
}

extern "C" int cf_debug_pow(int n, const double *x, double y, double *out) {
    if (n <= 0 || !x || !out) return CF_ERR_ARG;
    double *dx = nullptr, *dout = nullptr;
    HIPCHK(hipMalloc(&dx, sizeof(double) * (size_t)n));
    HIPCHK(hipMalloc(&dout, sizeof(double) * (size_t)n));
    HIPCHK(hipMemcpy(dx, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(pow_kernel, dim3(grid_of(n, 256)), dim3(256), 0, 0, dx, y, n, dout);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dout, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    dfree(dx);
    dfree(dout);
    return CF_OK;
}
