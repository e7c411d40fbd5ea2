// Schedule autotuner of a batch (cf_batch_autotune in cuteflow_hip.h).
//
// The step's speed depends on the cells per workgroup (256, 128 or 64: the
// share of a tile's cells lying on its edge, the waves per workgroup for
// the dt reduction, the grain of dry-tile skipping), on whether dry tiles
// are skipped (a stamp check per tile against the work it saves) and on
// the workgroup -> tile order (round-robin over the XCDs' separate L2s, or
// one contiguous run of tiles per XCD).  Which wins depends on the meshes,
// the wet fraction and the batch size, so it is measured: one handle per
// tile size, the other two knobs switched on that handle, each candidate
// timed over nsteps steps from the caller's initial states.
#include <chrono>
#include <vector>

#include "cf_internal.h"

using namespace cf;

namespace {

constexpr int kTuned = CF_OPT_TILE_128 | CF_OPT_TILE_64 | CF_OPT_NO_TILE_SKIP | CF_OPT_XCD_CONTIG;

int candidate_options(int base, int k) {
    int o = base & ~kTuned;
    if (k / 4 == 1) o |= CF_OPT_TILE_128;
    if (k / 4 == 2) o |= CF_OPT_TILE_64;
    if (k & 2) o |= CF_OPT_NO_TILE_SKIP;
    if (k & 1) o |= CF_OPT_XCD_CONTIG;
    return o;
}

// One run of nsteps from the initial states; wall time of the run alone
// (the step loop is one graph launch, ended by the status read-back).
int timed_run(cf_handle *h, const double *const *states, int nsteps, double &ms) {
    for (int b = 0; b < h->ninst; b++) {
        const int st = cf_batch_set_state(h, b, states[b]);
        if (st != CF_OK) return st;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    const auto t0 = std::chrono::steady_clock::now();
    const int st = cf_batch_step(h, nsteps, nullptr, nullptr, nullptr);
    const auto t1 = std::chrono::steady_clock::now();
    ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return st;
}

}  // namespace

extern "C" int cf_batch_autotune(int ninst, const cf_mesh *meshes, const cf_params *params, const double *const *states,
                                 int nsteps, int *best_options, double *times_ms) {
    if (ninst <= 0 || !meshes || !params || !states || !best_options || nsteps <= 0) return CF_ERR_ARG;
    for (int b = 0; b < ninst; b++)
        if (!states[b]) return CF_ERR_ARG;
    double t[CF_TUNE_NCAND];
    for (int k = 0; k < CF_TUNE_NCAND; k++) t[k] = -1.0;
    int first_err = CF_OK;
    std::vector<cf_params> prm(params, params + ninst);
    for (int ts = 0; ts < 3; ts++) {
        for (auto &p : prm) p.options = candidate_options(params[0].options, 4 * ts);
        cf_handle *h = nullptr;
        int st = cf_batch_create(ninst, meshes, prm.data(), &h);
        if (st != CF_OK) {
            if (first_err == CF_OK) first_err = st;
            continue;
        }
        for (int k = 4 * ts; k < 4 * ts + 4 && st == CF_OK; k++) {
            st = set_schedule(h, (k & 2) == 0, k & 1);
            double ms = 0.0, best = 1e300;
            // the first run captures the step graph and warms the caches
            for (int rep = 0; rep < 3 && st == CF_OK; rep++) {
                st = timed_run(h, states, nsteps, ms);
                if (rep > 0) best = std::min(best, ms);
            }
            if (st == CF_OK) t[k] = best;
            else if (first_err == CF_OK) first_err = st;
        }
        cf_destroy(h);
    }
    int kb = -1;
    for (int k = 0; k < CF_TUNE_NCAND; k++)
        if (t[k] >= 0.0 && (kb < 0 || t[k] < t[kb])) kb = k;
    // Near-ties (within 3 %, the spread of repeated short timings) go to the
    // fastest candidate that skips dry tiles: the skip also saves the replay
    // and the reverse sweep of derivative runs, which are not timed here.
    if (kb >= 0 && (kb & 2)) {
        int ks = -1;
        for (int k = 0; k < CF_TUNE_NCAND; k++)
            if (t[k] >= 0.0 && !(k & 2) && t[k] <= 1.03 * t[kb] && (ks < 0 || t[k] < t[ks])) ks = k;
        if (ks >= 0) kb = ks;
    }
    if (times_ms)
        for (int k = 0; k < CF_TUNE_NCAND; k++) times_ms[k] = t[k];
    if (kb < 0) return first_err != CF_OK ? first_err : CF_ERR_HIP;
    *best_options = candidate_options(params[0].options, kb);
    return CF_OK;
}
