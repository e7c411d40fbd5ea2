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
//     workgroup LDS pass to one plain store per workgroup; a one-workgroup
//     finalize kernel reduces those, publishes dt, tc and the stop/error
//     status for the next step, so a run of steps needs no host
//     synchronisation and no device-wide fence or atomic (on gfx950 an
//     agent-scope fence writes back and invalidates the XCD's L2);
//   - cells are stored along a Hilbert curve (256-cell tiles are compact
//     patches) and tiles that are dry with dry neighbours are skipped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/cuteflow_hip.h"
#include "cf_device.h"
#include "cf_prep.h"

#pragma clang fp contract(off)

using namespace cf;

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// geometry arrays (structure of arrays, each nelt doubles)
enum { G_N1 = 0, G_N2 = 3, G_LN = 6, G_SURF = 9, G_COT = 10, G_XC = 11, G_YC = 12, G_COUNT = 13 };

struct KParams {
    double gp, tolisec, tol, cfl, ts, constant_dt, override_manning;
    double pad_dt;  // min(1e6, 1e5 if nelt % 1024 != 0): reduction padding
    int ilimiteur, inlet, outlet, is_dt_constant, is_override_manning;
    int reserved;
    double debitglob[CF_MAX_IO];
    double h_sortie[CF_MAX_IO];
    double long_entree[CF_MAX_IO];
    int hydro_n[CF_MAX_IO];
};

struct Ctrl {
    int *hdr;          // [0] status (0 running, 1 ts reached, -1 dt/NaN failure), [1] steps done
    double *partial;   // per workgroup: min CFL dt of the new state, NaN if a state is NaN
    double *dt;        // dt[k] used by step k
    double *tc;        // tc[k] at the start of step k
    int *ran;          // ran[k]: tiles that ran step k (skipped tiles store +inf partials)
};

struct StepArgs {
    const double4 *__restrict__ st_in;
    double4 *__restrict__ st_out;
    const int4 *__restrict__ nbr;
    const double *__restrict__ geo;
    const double4 *__restrict__ dfc;
    const double *__restrict__ manning;
    const double *__restrict__ hydro;  // t[20*64] then q[20*64]
    Ctrl c;
    // dry-tile skipping: a tile runs step s only if it or a neighbour tile
    // held a wet cell after step s-1 or s-2 (stamp >= s-1), or always[tile]
    unsigned *stamp;
    const int *__restrict__ tile_off;
    const int *__restrict__ tile_adj;
    const unsigned char *__restrict__ always;
    unsigned gstep;  // global number of step 0 of this launch
    int skip;        // 1: dry-tile skipping enabled
    int nelt;
    int step;
    int xcd_map;  // 1: consecutive tiles on the same XCD
    int reserved;
    KParams p;
};

__device__ double inflow_q(const KParams &p, const double *hydro, int id, double tc) {
    int n = p.hydro_n[id];
    if (n <= 0) return p.debitglob[id];
    const double *t = hydro + id * CF_MAX_HYDRO;
    const double *q = hydro + CF_MAX_IO * CF_MAX_HYDRO + id * CF_MAX_HYDRO;
    if (tc <= t[0]) return q[0];
    for (int i = 0; i + 1 < n; i++)
        if (tc < t[i + 1]) return q[i] + (q[i + 1] - q[i]) * ((tc - t[i]) / (t[i + 1] - t[i]));
    return q[n - 1];
}

__device__ __forceinline__ int nb_at(const int4 &v, int j) { return j == 0 ? v.x : (j == 1 ? v.y : v.z); }
__device__ __forceinline__ double d4_at(const double4 &v, int j) { return j == 0 ? v.x : (j == 1 ? v.y : v.z); }

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = dmin(v, __shfl_xor(v, off, 64));
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap the workgroup
// id so that each XCD works on one contiguous range of tiles, whose
// neighbour gathers then hit that XCD's own L2.
__device__ __forceinline__ int tile_of(int b, int nblocks, int xcd_map) {
    if (!xcd_map) return b;
    const int q = nblocks >> 3, r = nblocks & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// Flux of a face whose two reconstructed depths are both below eps: the
// Riemann solvers return a zero flux there; same operations as their
// rotation back to the global frame.
__device__ __forceinline__ void zero_flux(double n1, double n2, double flux[3]) {
    const double z = 0.0;
    flux[0] = z;
    flux[1] = n1 * z - n2 * z;
    flux[2] = n2 * z + n1 * z;
}

// Occupancy target of the step kernel (waves per SIMD): 4 waves caps the
// kernel at 128 VGPRs without spilling (measured best on MI355X; 6 and 8
// spill); 0 lets the compiler choose the register budget.
#ifndef CF_STEP_WAVES_PER_EU
#define CF_STEP_WAVES_PER_EU 4
#endif
#if CF_STEP_WAVES_PER_EU > 0
#define CF_STEP_OCCUPANCY __attribute__((amdgpu_waves_per_eu(CF_STEP_WAVES_PER_EU, 8)))
#else
#define CF_STEP_OCCUPANCY
#endif

// FRIC: 0 no friction, 1 explicit (fricimplic 0), 2 fricimplic 1, 3 fricimplic 2
template <int IFLUX, int FRIC>
__global__ __launch_bounds__(kBlock) CF_STEP_OCCUPANCY void step_kernel(StepArgs a) {
    __shared__ double s_min[kWaves];
    if (a.c.hdr[0] != 0) return;  // stopped (ts reached) or failed earlier
    const int nelt = a.nelt;
    const int tile = tile_of(blockIdx.x, gridDim.x, a.xcd_map);
    const unsigned s_glob = a.gstep + (unsigned)a.step;
    // A tile whose cells and neighbour tiles' cells were all dry after the
    // last two steps would recompute (0, 0, 0) for every cell (all faces
    // dry/dry, no source, clamp), and both state buffers already hold that:
    // it is skipped.  Reading a stamp raised concurrently only makes the
    // tile run, which is always correct.
    if (a.skip && !a.always[tile] && a.stamp[tile] + 1u < s_glob) {
        if (threadIdx.x == 0) a.c.partial[blockIdx.x] = __builtin_inf();  // no dt contribution
        return;
    }
    const int i = tile * kBlock + threadIdx.x;
    int wet_local = 0;
    const double dt = a.c.dt[a.step];
    const double tc = a.c.tc[a.step];
    const KParams &p = a.p;
    const double g = p.gp, tls = p.tolisec;
    double dti = 1000000.0;
    int isnan_local = 0;

    if (i < nelt) {
        const double4 u0 = a.st_in[i];
        const int4 nb = a.nbr[i];
        const double zml = u0.w;
        // first ring: the three neighbour states, gathered once
        double4 unb[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int c = nb_at(nb, j);
            unb[j] = c >= 0 ? a.st_in[c] : u0;
        }
        double4 dfi = make_double4(0.0, 0.0, 0.0, 0.0);
        if (IFLUX == 3) dfi = a.dfc[i];
        double fl[3][3], fn1[3], fn2[3], fln[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double n1 = a.geo[(G_N1 + j) * nelt + i];
            const double n2 = a.geo[(G_N2 + j) * nelt + i];
            const double ln = a.geo[(G_LN + j) * nelt + i];
            fn1[j] = n1; fn2[j] = n2; fln[j] = ln;
            const int code = nb_at(nb, j);
            double hautl, vitul, vitvl;
            rot(tls, n1, n2, u0.x, u0.y, u0.z, hautl, vitul, vitvl);
            double flux[3] = {0.0, 0.0, 0.0};
            int kv = -1;
            if (code >= 0) {
                kv = code;
            } else {
                const int cls = (-code) & 3, id = (-code) >> 2;
                if (cls == 1) {  // inlet (flux_riemann.cuf:115-139)
                    if (p.inlet == 1) {
                        flux[0] = hautl * vitul;
                        flux[1] = n1 * (hautl * (vitul * vitul) + (g * (hautl * hautl)) / 2.0) - ((n2 * hautl) * vitul) * vitvl;
                        flux[2] = n2 * (hautl * (vitul * vitul) + (g * (hautl * hautl)) / 2.0) + ((n1 * hautl) * vitul) * vitvl;
                    } else if (p.inlet == 2) {
                        const double dn = -(inflow_q(p, a.hydro, id - 1, tc) / p.long_entree[id - 1]);
                        flux[0] = dn;
                        flux[1] = n1 * ((dn * dn) / (hautl + tls) + (g * (hautl * hautl)) / 2.0);
                        flux[2] = n2 * ((dn * dn) / (hautl + tls) + (g * (hautl * hautl)) / 2.0);
                    }
                } else if (cls == 3) {  // wall (:141-145)
                    flux[1] = ((n1 * g) * (hautl * hautl)) / 2.0;
                    flux[2] = ((n2 * g) * (hautl * hautl)) / 2.0;
                } else if (cls == 2) {
                    if (p.outlet == 1) {  // fixed level (:147-163)
                        if (vitul >= 0.0 && vitul / sqrt(g * hautl) < 1.0) {
                            hautl = dmax(0.0, p.h_sortie[id - 1] - zml);
                            flux[0] = hautl * vitul;
                            flux[1] = n1 * (hautl * (vitul * vitul) + (g * (hautl * hautl)) / 2.0) - ((n2 * hautl) * vitul) * vitvl;
                            flux[2] = n2 * (hautl * (vitul * vitul) + (g * (hautl * hautl)) / 2.0) + ((n1 * hautl) * vitul) * vitvl;
                        } else {
                            flux[1] = ((n1 * g) * (hautl * hautl)) / 2.0;
                            flux[2] = ((n2 * g) * (hautl * hautl)) / 2.0;
                        }
                    } else {  // transmissive: Riemann problem with itself (:166-171)
                        kv = i;
                    }
                }
            }
            if (kv >= 0) {  // interior face (:166-409)
                const double4 ur4 = (kv == i) ? u0 : unb[j];
                const double zmr = ur4.w;
                const double zij = dmax(zml, zmr);
                double hijet = dmax(0.0, (hautl + zml) - zij);
                double hautr, vitur, vitvr;
                rot(tls, n1, n2, ur4.x, ur4.y, ur4.z, hautr, vitur, vitvr);
                if (IFLUX == 1) {
                    double hjiet = dmax(0.0, (ur4.x + zmr) - zij);
                    if (hijet < tls) { hijet = 0.0; vitul = 0.0; vitvl = 0.0; }
                    if (hjiet < tls) { hjiet = 0.0; vitur = 0.0; vitvr = 0.0; }
                    hllc_flux(g, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr, flux);
                } else {
                    // reconstructed depths of flux_riemann.cuf:371-394 (the
                    // values solver_Riemann_waf receives)
                    double hjiet = dmax(0.0, (hautr + zmr) - zij);
                    if (hijet < tls) { hijet = 0.0; vitul = 0.0; vitvl = 0.0; }
                    if (hjiet < tls) { hjiet = 0.0; vitur = 0.0; vitvr = 0.0; }
                    if (hijet < kEpsWaf && hjiet < kEpsWaf) {
                        zero_flux(n1, n2, flux);  // dry/dry: no upwind data needed
                    } else {
                        const double d_mrl = d4_at(dfi, j);
                        const int ja = (j + 1) % 3, jb = (j + 2) % 3;
                        const bool own1 = nb_at(nb, ja) < 0, own2 = nb_at(nb, jb) < 0;
                        const double4 s1 = own1 ? u0 : unb[ja];
                        const double4 s2 = own2 ? u0 : unb[jb];
                        double h1, u1, w1, h2, u2, w2;
                        rot(tls, n1, n2, s1.x, s1.y, s1.z, h1, u1, w1);
                        rot(tls, n1, n2, s2.x, s2.y, s2.z, h2, u2, w2);
                        const int tjv = (nb.w >> (2 * j)) & 3;
                        const int4 nbk = (kv == i) ? nb : a.nbr[kv];
                        const int ka = (tjv + 1) % 3, kb = (tjv + 2) % 3;
                        const int e1 = nb_at(nbk, ka), e2 = nb_at(nbk, kb);
                        const int kw1 = e1 >= 0 ? e1 : kv, kw2 = e2 >= 0 ? e2 : kv;
                        const double4 g1 = a.st_in[kw1], g2 = a.st_in[kw2];
                        double g1h, g1u, g1w, g2h, g2u, g2w;
                        rot(tls, n1, n2, g1.x, g1.y, g1.z, g1h, g1u, g1w);
                        rot(tls, n1, n2, g2.x, g2.y, g2.z, g2h, g2u, g2w);
                        double hl_upw, vl_upw, hr_upw, vr_upw, zml_upw, zmr_upw, d_mkvol, d_mkvoisin;
                        if ((own1 && own2) || (kw1 == kv && kw2 == kv)) {
                            hl_upw = hautl; vl_upw = vitvl;
                            hr_upw = hautr; vr_upw = vitvr;
                            d_mkvol = d_mrl; d_mkvoisin = d_mrl;
                            zml_upw = zml; zmr_upw = zmr;
                        } else {
                            if (own1) {
                                hl_upw = h2; vl_upw = w2;
                                d_mkvol = d4_at(dfi, jb);
                                zml_upw = s2.w;
                            } else if (own2) {
                                hl_upw = h1; vl_upw = w1;
                                d_mkvol = d4_at(dfi, ja);
                                zml_upw = s1.w;
                            } else {
                                const double d1 = d4_at(dfi, ja), d2 = d4_at(dfi, jb);
                                hl_upw = (d1 * h1 + d2 * h2) / (d1 + d2);
                                vl_upw = (d1 * w1 + d2 * w2) / (d1 + d2);
                                zml_upw = (d1 * s1.w + d2 * s2.w) / (d1 + d2);
                                d_mkvol = (d1 + d2) / 2.0;
                            }
                            const double4 dfk = (kv == i) ? dfi : a.dfc[kv];
                            if (kw1 == kv) {
                                hr_upw = g2h; vr_upw = g2w;
                                d_mkvoisin = d4_at(dfk, kb);
                                zmr_upw = g2.w;
                            } else if (kw2 == kv) {
                                hr_upw = g1h; vr_upw = g1w;
                                d_mkvoisin = d4_at(dfk, ka);
                                zmr_upw = g1.w;
                            } else {
                                const double d1 = d4_at(dfk, ka), d2 = d4_at(dfk, kb);
                                hr_upw = (d1 * g1h + d2 * g2h) / (d1 + d2);
                                vr_upw = (d1 * g1w + d2 * g2w) / (d1 + d2);
                                zmr_upw = (d1 * g1.w + d2 * g2.w) / (d1 + d2);
                                d_mkvoisin = (d1 + d2) / 2.0;
                            }
                        }
                        // upwind hydrostatic reconstruction (flux_riemann.cuf:373,376,384-400)
                        hl_upw = dmax(0.0, (hl_upw + zml_upw) - zij);
                        hr_upw = dmax(0.0, (hr_upw + zmr_upw) - zij);
                        if (hl_upw < tls) { hl_upw = 0.0; vl_upw = 0.0; }
                        if (hr_upw < tls) { hr_upw = 0.0; vr_upw = 0.0; }
                        waf_flux(g, p.ilimiteur, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr,
                                 hl_upw, vl_upw, hr_upw, vr_upw, d_mrl, d_mkvol, d_mkvoisin, dt, flux);
                    }
                }
            }
            fl[j][0] = flux[0] * ln;
            fl[j][1] = flux[1] * ln;
            fl[j][2] = flux[2] * ln;
        }
        // face sum (flux_riemann.cuf:411-426)
        const double resc0 = (fl[0][0] + fl[1][0]) + fl[2][0];
        const double resc1 = (fl[0][1] + fl[1][1]) + fl[2][1];
        const double resc2 = (fl[0][2] + fl[1][2]) + fl[2][2];

        // bed slope source over interior faces (source_terms.cuf:50-100)
        const double q1 = u0.x, q2 = u0.y, q3 = u0.z;
        double sb1 = 0.0, sb2 = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (nb_at(nb, j) >= 0) {
                const double zij = dmax(zml, unb[j].w);
                const double hijet = dmax(0.0, (q1 + zml) - zij);
                sb1 = sb1 + (((g / 2.0) * (hijet * hijet - q1 * q1)) * fln[j]) * fn1[j];
                sb2 = sb2 + (((g / 2.0) * (hijet * hijet - q1 * q1)) * fln[j]) * fn2[j];
            }
        }
        const double surf = a.geo[G_SURF * nelt + i];
        double r0, r1, r2;
        if (FRIC == 0) {
            // update without friction (finite_volumes_method.cuf:294)
            r0 = (0.0 - resc0) / surf;
            r1 = (sb1 - resc1) / surf;
            r2 = (sb2 - resc2) / surf;
        } else {
            // friction (source_terms.cuf:102-209)
            const double tol = p.tol;
            const double mann = p.is_override_manning ? p.override_manning : a.manning[i];
            double sf1 = 0.0, sf2 = 0.0, vitu = 0.0, vitv = 0.0, normvit = 0.0;
            if (q1 > tls) {
                vitu = q2 / q1;
                vitv = q3 / q1;
                normvit = sqrt(vitu * vitu + vitv * vitv);
                if (normvit < tol) normvit = tol;
                const double p13 = cf_pow(q1, -1.0 / 3.0);
                sf1 = -(((((g * mann) * mann) * vitu) * normvit) * surf * p13);
                sf2 = -(((((g * mann) * mann) * vitv) * normvit) * surf * p13);
            }
            const double x0 = ((0.0 + 0.0) - resc0) / surf;
            const double x1 = ((sb1 + sf1) - resc1) / surf;
            const double x2 = ((sb2 + sf2) - resc2) / surf;
            double a0 = 1.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 1.0, a5 = 0.0, a6 = 0.0, a7 = 0.0, a8 = 1.0;
            if (FRIC == 2 && normvit > tol && q1 > tls) {
                const double dtgn2h = ((((dt / 2.0) * g) * mann) * mann) * cf_pow(q1, -4.0 / 3.0);
                const double f21 = -((((7.0 / 3.0) * dtgn2h) * vitu) * normvit);
                const double f22 = 1.0 + (dtgn2h * (2.0 * (vitu * vitu) + vitv * vitv)) / normvit;
                const double f23 = ((dtgn2h * vitu) * vitv) / normvit;
                const double f31 = -((((7.0 / 3.0) * dtgn2h) * vitv) * normvit);
                const double f32 = f23;
                const double f33 = 1.0 + (dtgn2h * (vitu * vitu + 2.0 * (vitv * vitv))) / normvit;
                const double det = f22 * f33 - f23 * f32;
                a3 = -(f21 * f33 - f23 * f31) / det;
                a4 = f33 / det;
                a5 = -f23 / det;
                a6 = (f21 * f32 - f22 * f31) / det;
                a7 = -f32 / det;
                a8 = f22 / det;
            } else if (FRIC == 3 && normvit > tol && q1 > tls) {
                const double beta = -(((((dt * g) * mann) * mann) * normvit) * cf_pow(q1, -4.0 / 3.0));
                const double det = (1.0 - beta) * (1.0 - beta);
                a4 = (1.0 - beta) / det;
                a8 = (1.0 - beta) / det;
            }
            if (FRIC == 1) {
                r0 = x0; r1 = x1; r2 = x2;
            } else {
                // res = afm1 * resini (finite_volumes_method.cuf:286-289)
                r0 = (a0 * x0 + a1 * x1) + a2 * x2;
                r1 = (a3 * x0 + a4 * x1) + a5 * x2;
                r2 = (a6 * x0 + a7 * x1) + a8 * x2;
            }
        }
        // update and wet/dry clamp (finite_volumes_method.cuf:297-303)
        double h = u0.x + dt * r0;
        double hu = u0.y + dt * r1;
        double hv = u0.z + dt * r2;
        if (h <= tls) { h = 0.0; hu = 0.0; hv = 0.0; }
        a.st_out[i] = make_double4(h, hu, hv, zml);
        isnan_local = (h != h) || (hu != hu) || (hv != hv);
        wet_local = (h != 0.0) || isnan_local;
        // cflcond on the new state (cfl_condition.cuf:20-36)
        if (h > tls) {
            const double uu = hu / h, vv = hv / h;
            const double nv = sqrt(uu * uu + vv * vv);
            const double vmax = dmax(0.01, nv + sqrt(g * fabs(h)));
            dti = dmin((p.cfl * a.geo[G_COT * nelt + i]) / vmax, 1000000.0);
        }
        if (p.is_dt_constant == 1) dti = p.constant_dt;
    }

    // per-workgroup dt minimum: wavefront shuffles, then LDS across waves;
    // one plain store per workgroup (reduced by finalize_kernel)
    const double m = wave_min(dti);
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = m;
    const int anynan = __syncthreads_or(isnan_local);
    if (threadIdx.x == 0) {
        double bm = s_min[0];
#pragma unroll
        for (int w = 1; w < kWaves; w++) bm = dmin(bm, s_min[w]);
        a.c.partial[blockIdx.x] = anynan ? __builtin_nan("") : bm;
    }
    // a wet tile keeps itself and its neighbour tiles running for two steps
    if (a.skip && __syncthreads_or(wet_local)) {
        const int o = a.tile_off[tile], n = a.tile_off[tile + 1] - o;
        for (int k = threadIdx.x; k <= n; k += kBlock)
            a.stamp[k == 0 ? tile : a.tile_adj[o + k - 1]] = s_glob + 1u;
    }
}

// Reduces a set of per-workgroup minima (NaN marks a NaN state) starting
// from `pad`; thread 0 gets the result and the NaN flag.
__device__ __forceinline__ double block_min_1024(const double *__restrict__ part, int n, double pad, int &nan,
                                                 int &nskip) {
    __shared__ double s[1024 / 64];
    __shared__ int sk[1024 / 64];
    double m = pad;
    int bad = 0, skipped = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double v = part[k];
        if (v != v) bad = 1;
        else m = dmin(m, v);
        skipped += (v == __builtin_inf());
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

// End of a step: dt of the new state (the reference's reduction, cfl_condition.cuf:68-104),
// tc += dt, dt sanity check (:129-132) and the loop condition (finite_volumes_method.cuf:123).
__global__ __launch_bounds__(1024) void finalize_kernel(Ctrl c, int nparts, int step, double pad,
                                                         double ts, double tol) {
    if (c.hdr[0] != 0) return;
    int nan, nskip;
    const double dtn = block_min_1024(c.partial, nparts, pad, nan, nskip);
    if (threadIdx.x == 0) {
        const double tcn = c.tc[step] + c.dt[step];
        c.dt[step + 1] = dtn;
        c.tc[step + 1] = tcn;
        c.ran[step] = nparts - nskip;
        c.hdr[1] = step + 1;
        if (nan || !(dtn <= 9999.0 && dtn >= kDtLow)) c.hdr[0] = -1;
        else if (!(((tcn - ts) + 2.0 * tol) < tol)) c.hdr[0] = 1;
    }
}

// dt of the current state (finite_volumes_method.cuf:113-116), into *out.
__global__ __launch_bounds__(1024) void reduce_dt_kernel(const double *part, int nparts, double pad, double *out) {
    int nan, nskip;
    const double m = block_min_1024(part, nparts, pad, nan, nskip);
    if (threadIdx.x == 0) *out = nan ? __builtin_nan("") : m;
}

// Prepares the control block for a run of steps.
__global__ void ctrl_init_kernel(Ctrl c, double dt0, double tc0, double ts, double tol) {
    c.dt[0] = dt0;
    c.tc[0] = tc0;
    c.hdr[1] = 0;
    c.hdr[0] = (((tc0 - ts) + 2.0 * tol) < tol) ? 0 : 1;
}

// cflcond for the current state, one minimum per workgroup (cfl_condition.cuf:4-39);
// like the reference, the initial dt does not look for NaN states
__global__ __launch_bounds__(kBlock) void cfl_kernel(const double4 *__restrict__ st,
                                                      const double *__restrict__ cot, int nelt,
                                                      KParams p, double *__restrict__ part) {
    __shared__ double s_min[kWaves];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double dti = 1000000.0;
    if (i < nelt) {
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
        for (int w = 1; w < kWaves; w++) bm = dmin(bm, s_min[w]);
        part[blockIdx.x] = bm;
    }
}


// Geometry of each cell (maillage.cuf:132-159, :1254-1331, flux_riemann.cuf:50-74).
__global__ void geometry_kernel(const int *__restrict__ conn, const double *__restrict__ coords,
                                int nelt, double *__restrict__ geo, double *__restrict__ zm_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nelt) return;
    const int s1 = conn[3 * i] - 1, s2 = conn[3 * i + 1] - 1, s3 = conn[3 * i + 2] - 1;
    const double vx[3] = {coords[3 * s1], coords[3 * s2], coords[3 * s3]};
    const double vy[3] = {coords[3 * s1 + 1], coords[3 * s2 + 1], coords[3 * s3 + 1]};
    const double vz[3] = {coords[3 * s1 + 2], coords[3 * s2 + 2], coords[3 * s3 + 2]};
    const double x1 = vx[0], y1 = vy[0], x2 = vx[1], y2 = vy[1], x3 = vx[2], y3 = vy[2];
    geo[G_XC * nelt + i] = ((x1 + x2) + x3) / 3.0;
    geo[G_YC * nelt + i] = ((y1 + y2) + y3) / 3.0;
    {
        const double ca = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
        const double cb = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3));
        const double cc = sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1));
        const double pp = ((ca + cb) + cc) / 2.0;
        const double aire = sqrt(((pp * (pp - ca)) * (pp - cb)) * (pp - cc));
        double cot = (0.5 * aire) / pp;
        geo[G_COT * nelt + i] = kCote * cot;
    }
    {
        const double aa = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
        const double bb = sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
        const double cc = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3));
        const double s = ((aa + bb) + cc) / 2.0;
        geo[G_SURF * nelt + i] = sqrt(((s * (s - aa)) * (s - bb)) * (s - cc));
        const double xg = ((x1 + x2) + x3) / 3.0, yg = ((y1 + y2) + y3) / 3.0;
        const double d1 = sqrt((x1 - xg) * (x1 - xg) + (y1 - yg) * (y1 - yg));
        const double d2 = sqrt((x2 - xg) * (x2 - xg) + (y2 - yg) * (y2 - yg));
        const double d3 = sqrt((x3 - xg) * (x3 - xg) + (y3 - yg) * (y3 - yg));
        zm_out[i] = ((d1 * vz[0] + d2 * vz[1]) + d3 * vz[2]) / ((d1 + d2) + d3);
    }
    for (int j = 0; j < 3; j++) {
        const int c1 = j, c2 = (j + 1) % 3, c3 = (j + 2) % 3;
        const double ax = vx[c1], ay = vy[c1], bx = vx[c2], by = vy[c2], cx = vx[c3], cy = vy[c3];
        const double l = sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay));
        double n1 = (by - ay) / l, n2 = (ax - bx) / l;
        const double orient = n1 * ((ax + bx) / 2.0 - ((ax + bx) + cx) / 3.0) +
                              n2 * ((ay + by) / 2.0 - ((ay + by) + cy) / 3.0);
        if (orient < 0.0) {
            n1 = (ay - by) / l;
            n2 = (bx - ax) / l;
        }
        geo[(G_N1 + j) * nelt + i] = n1;
        geo[(G_N2 + j) * nelt + i] = n2;
        geo[(G_LN + j) * nelt + i] = l;
    }
}

// Centroid distance across each face (0 across a boundary).
__global__ void face_dist_kernel(const int4 *__restrict__ nbr, const double *__restrict__ geo,
                                 int nelt, double4 *__restrict__ dfc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nelt) return;
    const int4 nb = nbr[i];
    const double xi = geo[G_XC * nelt + i], yi = geo[G_YC * nelt + i];
    double d[3];
    for (int j = 0; j < 3; j++) {
        const int k = nb_at(nb, j);
        const int kk = k >= 0 ? k : i;
        const double xk = geo[G_XC * nelt + kk], yk = geo[G_YC * nelt + kk];
        d[j] = sqrt((xi - xk) * (xi - xk) + (yi - yk) * (yi - yk));
    }
    dfc[i] = make_double4(d[0], d[1], d[2], 0.0);
}

// State in the caller's order and layout (3, nelt) <-> internal double4
// (h, hu, hv, zm) in the internal cell order; perm: internal -> caller.
__global__ void pack_state_kernel(const double *__restrict__ U, const double *__restrict__ zm,
                                  const int *__restrict__ perm, int nelt, double4 *__restrict__ st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nelt) {
        const size_t o = 3 * (size_t)perm[i];
        st[i] = make_double4(U[o], U[o + 1], U[o + 2], zm[i]);
    }
}

__global__ void unpack_state_kernel(const double4 *__restrict__ st, const int *__restrict__ perm,
                                    int nelt, double *__restrict__ U) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nelt) {
        const double4 s = st[i];
        const size_t o = 3 * (size_t)perm[i];
        U[o] = s.x;
        U[o + 1] = s.y;
        U[o + 2] = s.z;
    }
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

struct cf_handle {
    int nelt = 0, nnt = 0;
    cf_params prm{};
    KParams kp{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double4 *st[2] = {nullptr, nullptr};
    int cur = 0;
    int4 *nbr = nullptr;
    double *geo = nullptr;
    double4 *dfc = nullptr;
    double *manning = nullptr;
    double *hydro = nullptr;
    double *zm = nullptr;
    double *stage = nullptr;
    Ctrl c{};
    int cap = 0;
    double *dt0 = nullptr;     // initial dt of the current state
    int nblocks = 0;           // workgroups of a step launch (one tile each)
    int xcd_map = 0;
    // internal cell order (internal -> caller) on the device and the host
    int *perm = nullptr;
    std::vector<int> hperm;
    // dry-tile skipping
    unsigned *stamp = nullptr;
    int *tile_off = nullptr;
    int *tile_adj = nullptr;
    unsigned char *always = nullptr;
    unsigned gstep = 0;        // steps run since creation
    int skip = 1;
    long long last_tiles_run = 0, last_tiles_total = 0;
    double dt = 0.0, tc = 0.0;
    int status = 0;
    int fric = 0;
    int pending = -1;  // steps enqueued by cf_step_async, not yet finished
};

#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) {                     \
            fprintf(stderr, "cuteflow_hip: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return CF_ERR_HIP;                      \
        }                                           \
    } while (0)

static int grid_of(int n, int b) { return (n + b - 1) / b; }

template <typename T>
static void dfree(T *&ptr) {
    if (ptr) (void)hipFree((void *)ptr);
    ptr = nullptr;
}

static void fill_kparams(cf_handle *h) {
    const cf_params &p = h->prm;
    KParams &k = h->kp;
    k.gp = p.gp; k.tolisec = p.tolisec; k.tol = p.tol; k.cfl = p.cfl; k.ts = p.ts;
    k.constant_dt = p.constant_dt; k.override_manning = p.override_manning;
    k.pad_dt = (h->nelt % 1024 != 0) ? 100000.0 : 1000000.0;
    k.ilimiteur = p.ilimiteur; k.inlet = p.inlet; k.outlet = p.outlet;
    k.is_dt_constant = p.is_dt_constant; k.is_override_manning = p.is_override_manning;
    for (int i = 0; i < CF_MAX_IO; i++) {
        k.debitglob[i] = p.debitglob[i];
        k.h_sortie[i] = p.h_sortie[i];
        k.long_entree[i] = p.long_entree[i];
        k.hydro_n[i] = p.hydro_n[i];
    }
}

static int upload_hydro(cf_handle *h) {
    std::vector<double> buf(2 * CF_MAX_IO * CF_MAX_HYDRO);
    memcpy(buf.data(), h->prm.hydro_t, sizeof(double) * CF_MAX_IO * CF_MAX_HYDRO);
    memcpy(buf.data() + CF_MAX_IO * CF_MAX_HYDRO, h->prm.hydro_q, sizeof(double) * CF_MAX_IO * CF_MAX_HYDRO);
    HIPCHK(hipMemcpyAsync(h->hydro, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

static int compute_dt(cf_handle *h) {
    hipLaunchKernelGGL(cfl_kernel, dim3(h->nblocks), dim3(kBlock), 0, h->stream,
                       h->st[h->cur], h->geo + G_COT * h->nelt, h->nelt, h->kp, h->c.partial);
    hipLaunchKernelGGL(reduce_dt_kernel, dim3(1), dim3(1024), 0, h->stream, h->c.partial, h->nblocks,
                       h->kp.pad_dt, h->dt0);
    HIPCHK(hipGetLastError());
    double dt = 0.0;
    HIPCHK(hipMemcpyAsync(&dt, h->dt0, sizeof(dt), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->dt = dt;
    return CF_OK;
}

static int ensure_ctrl(cf_handle *h, int nsteps) {
    if (nsteps <= h->cap) return CF_OK;
    int cap = std::max(nsteps, 2 * h->cap);
    if (h->c.dt) {
        HIPCHK(hipStreamSynchronize(h->stream));
        dfree(h->c.dt); dfree(h->c.tc); dfree(h->c.ran);
    }
    HIPCHK(hipMalloc(&h->c.dt, sizeof(double) * ((size_t)cap + 1)));
    HIPCHK(hipMalloc(&h->c.tc, sizeof(double) * ((size_t)cap + 1)));
    HIPCHK(hipMalloc(&h->c.ran, sizeof(int) * ((size_t)cap + 1)));
    h->cap = cap;
    return CF_OK;
}

extern "C" const char *cf_error_string(int s) {
    switch (s) {
    case CF_OK: return "ok";
    case CF_ERR_ARG: return "invalid argument or unsupported option";
    case CF_ERR_HIP: return "HIP runtime error";
    case CF_ERR_DT: return "time step out of range (dt > 9999, dt < 1e-12) or NaN state";
    case CF_ERR_MESH: return "inconsistent mesh";
    case CF_ERR_NOMEM: return "out of memory";
    default: return "unknown status";
    }
}

static int validate_params(const cf_params *p) {
    if (!(p->gp > 0.0) || !(p->cfl > 0.0) || !(p->tolisec >= 0.0)) return CF_ERR_ARG;
    if (p->iflux != 1 && p->iflux != 3) return CF_ERR_ARG;
    if (p->inlet != 1 && p->inlet != 2) return CF_ERR_ARG;  // 'phys_inflow' not provided
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
    dfree(h->st[0]); dfree(h->st[1]); dfree(h->nbr); dfree(h->geo); dfree(h->dfc);
    dfree(h->manning); dfree(h->hydro); dfree(h->zm); dfree(h->stage);
    dfree(h->c.hdr); dfree(h->c.partial); dfree(h->c.dt); dfree(h->c.tc); dfree(h->c.ran);
    dfree(h->dt0); dfree(h->perm); dfree(h->stamp); dfree(h->tile_off); dfree(h->tile_adj);
    dfree(h->always);
}

template <typename T>
static bool dalloc(T *&p, size_t count) {
    return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)) == hipSuccess;
}

template <typename T>
static bool h2d(T *dst, const T *src, size_t count, hipStream_t s) {
    return count == 0 || hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s) == hipSuccess;
}

extern "C" int cf_create(const cf_mesh *mesh, const cf_params *params, cf_handle **out) {
    if (!mesh || !params || !out || mesh->nelt <= 0 || mesh->nnt <= 0 || !mesh->conn ||
        !mesh->coords || !mesh->boundary)
        return CF_ERR_ARG;
    int st = validate_params(params);
    if (st) return st;
    *out = nullptr;
    const int nelt = mesh->nelt;
    const size_t N = (size_t)nelt;
    if (N > (size_t)INT32_MAX / 4) return CF_ERR_ARG;
    for (size_t i = 0; i < 3 * N; i++)
        if (mesh->conn[i] < 1 || mesh->conn[i] > mesh->nnt) return CF_ERR_MESH;

    cfprep::Prep prep;
    st = cfprep::prepare(mesh, params, kBlock, (params->options & CF_OPT_KEEP_ORDER) != 0, prep);
    if (st) return st;
    std::vector<int4> nbr(N);
    for (size_t i = 0; i < N; i++)
        nbr[i] = make_int4(prep.nbr[i].c[0], prep.nbr[i].c[1], prep.nbr[i].c[2], prep.nbr[i].w);
    std::vector<double> mann(N, 0.0);
    if (mesh->manning)
        for (size_t i = 0; i < N; i++) mann[i] = mesh->manning[prep.perm[i]];
    const int ntiles = (int)prep.tile_always.size();

    cf_handle *h = new (std::nothrow) cf_handle();
    if (!h) return CF_ERR_NOMEM;
    h->nelt = nelt;
    h->nnt = mesh->nnt;
    h->prm = *params;
    h->fric = params->friction == 0 ? 0 : 1 + params->fricimplic;
    h->nblocks = grid_of(nelt, kBlock);
    h->hperm = prep.perm;
    // skipping dry tiles is exact except with a constant dt (every cell then
    // contributes constant_dt to the minimum)
    h->skip = (params->options & CF_OPT_NO_TILE_SKIP) || params->is_dt_constant == 1 ? 0 : 1;
    fill_kparams(h);
    int rc = CF_OK;
    int *dconn = nullptr;
    double *dcoords = nullptr;
    do {
        if (hipStreamCreate(&h->stream) != hipSuccess) { rc = CF_ERR_HIP; break; }
        h->own_stream = true;
        if (!dalloc(h->st[0], N) || !dalloc(h->st[1], N) || !dalloc(h->nbr, N) ||
            !dalloc(h->geo, G_COUNT * N) || !dalloc(h->dfc, N) || !dalloc(h->manning, N) ||
            !dalloc(h->hydro, 2 * CF_MAX_IO * CF_MAX_HYDRO) || !dalloc(h->zm, N) ||
            !dalloc(h->stage, 3 * N) || !dalloc(h->c.hdr, 4) || !dalloc(h->dt0, 1) ||
            !dalloc(h->c.partial, (size_t)h->nblocks) || !dalloc(h->perm, N) ||
            !dalloc(h->stamp, (size_t)ntiles) || !dalloc(h->tile_off, prep.tile_off.size()) ||
            !dalloc(h->tile_adj, prep.tile_adj.size()) || !dalloc(h->always, (size_t)ntiles) ||
            !dalloc(dconn, 3 * N) || !dalloc(dcoords, 3 * (size_t)mesh->nnt)) {
            rc = CF_ERR_NOMEM;
            break;
        }
        if (!h2d(dconn, prep.conn.data(), 3 * N, h->stream) ||
            !h2d(dcoords, mesh->coords, 3 * (size_t)mesh->nnt, h->stream) ||
            !h2d(h->nbr, nbr.data(), N, h->stream) || !h2d(h->manning, mann.data(), N, h->stream) ||
            !h2d(h->perm, prep.perm.data(), N, h->stream) ||
            !h2d(h->tile_off, prep.tile_off.data(), prep.tile_off.size(), h->stream) ||
            !h2d(h->tile_adj, prep.tile_adj.data(), prep.tile_adj.size(), h->stream) ||
            !h2d(h->always, prep.tile_always.data(), (size_t)ntiles, h->stream) ||
            hipMemsetAsync(h->c.hdr, 0, 4 * sizeof(int), h->stream) != hipSuccess ||
            hipMemsetAsync(h->stamp, 0, (size_t)ntiles * sizeof(unsigned), h->stream) != hipSuccess ||
            hipMemsetAsync(h->stage, 0, 3 * N * sizeof(double), h->stream) != hipSuccess) {
            rc = CF_ERR_HIP;
            break;
        }
        hipLaunchKernelGGL(geometry_kernel, dim3(grid_of(nelt, 256)), dim3(256), 0, h->stream, dconn, dcoords, nelt, h->geo, h->zm);
        hipLaunchKernelGGL(face_dist_kernel, dim3(grid_of(nelt, 256)), dim3(256), 0, h->stream, h->nbr, h->geo, nelt, h->dfc);
        // zero state with the bed level in the 4th component, in both buffers
        hipLaunchKernelGGL(pack_state_kernel, dim3(grid_of(nelt, 256)), dim3(256), 0, h->stream, h->stage, h->zm, h->perm, nelt, h->st[0]);
        hipLaunchKernelGGL(pack_state_kernel, dim3(grid_of(nelt, 256)), dim3(256), 0, h->stream, h->stage, h->zm, h->perm, nelt, h->st[1]);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) { rc = CF_ERR_HIP; break; }
        if ((rc = upload_hydro(h)) != CF_OK) break;
        if ((rc = ensure_ctrl(h, 1024)) != CF_OK) break;
        if ((rc = compute_dt(h)) != CF_OK) break;
    } while (0);
    dfree(dconn);
    dfree(dcoords);
    if (rc != CF_OK) {
        destroy_buffers(h);
        if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return rc;
    }
    *out = h;
    return CF_OK;
}

extern "C" void cf_destroy(cf_handle *h) {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    destroy_buffers(h);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" int cf_set_state_device(cf_handle *h, const double *dU) {
    if (!h || !dU || h->pending >= 0) return CF_ERR_ARG;
    hipLaunchKernelGGL(pack_state_kernel, dim3(grid_of(h->nelt, 256)), dim3(256), 0, h->stream, dU, h->zm,
                       h->perm, h->nelt, h->st[h->cur]);
    // every tile runs the next two steps (both state buffers become consistent)
    const int ntiles = h->nblocks;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_of(ntiles, 256)), dim3(256), 0, h->stream, h->stamp, ntiles, h->gstep);
    HIPCHK(hipGetLastError());
    h->tc = 0.0;
    h->status = 0;
    return compute_dt(h);
}

extern "C" int cf_set_state(cf_handle *h, const double *U) {
    if (!h || !U || h->pending >= 0) return CF_ERR_ARG;
    HIPCHK(hipMemcpyAsync(h->stage, U, 3 * (size_t)h->nelt * sizeof(double), hipMemcpyHostToDevice, h->stream));
    return cf_set_state_device(h, h->stage);
}

extern "C" int cf_get_state_device(cf_handle *h, double *dU) {
    if (!h || !dU || h->pending >= 0) return CF_ERR_ARG;
    hipLaunchKernelGGL(unpack_state_kernel, dim3(grid_of(h->nelt, 256)), dim3(256), 0, h->stream, h->st[h->cur],
                       h->perm, h->nelt, dU);
    HIPCHK(hipGetLastError());
    return CF_OK;
}

extern "C" int cf_get_state(cf_handle *h, double *U) {
    if (!h || !U) return CF_ERR_ARG;
    int rc = cf_get_state_device(h, h->stage);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(U, h->stage, 3 * (size_t)h->nelt * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

extern "C" int cf_initial_dt(cf_handle *h, double *dt) {
    if (!h || !dt) return CF_ERR_ARG;
    *dt = h->dt;
    return CF_OK;
}

extern "C" int cf_get_time(cf_handle *h, double *tc, double *dt) {
    if (!h) return CF_ERR_ARG;
    if (tc) *tc = h->tc;
    if (dt) *dt = h->dt;
    return CF_OK;
}

template <int IFLUX, int FRIC>
static void launch_step(cf_handle *h, StepArgs &a, int grid) {
    hipLaunchKernelGGL((step_kernel<IFLUX, FRIC>), dim3(grid), dim3(kBlock), 0, h->stream, a);
}

typedef void (*step_launcher)(cf_handle *, StepArgs &, int);

static step_launcher pick_launcher(int iflux, int fric) {
    static const step_launcher tab[2][4] = {
        {launch_step<1, 0>, launch_step<1, 1>, launch_step<1, 2>, launch_step<1, 3>},
        {launch_step<3, 0>, launch_step<3, 1>, launch_step<3, 2>, launch_step<3, 3>}};
    return tab[iflux == 3 ? 1 : 0][fric];
}

extern "C" int cf_step_async(cf_handle *h, int nsteps) {
    if (!h || nsteps < 0) return CF_ERR_ARG;
    if (h->pending >= 0) return CF_ERR_ARG;  // previous run not finished
    if (h->status < 0) return CF_ERR_DT;
    int rc = ensure_ctrl(h, std::max(nsteps, 1));
    if (rc) return rc;
    h->pending = nsteps;
    if (nsteps == 0) return CF_OK;
    hipLaunchKernelGGL(ctrl_init_kernel, dim3(1), dim3(1), 0, h->stream, h->c, h->dt, h->tc, h->kp.ts,
                       h->kp.tol);
    StepArgs a;
    a.geo = h->geo;
    a.nbr = h->nbr;
    a.dfc = h->dfc;
    a.manning = h->manning;
    a.hydro = h->hydro;
    a.c = h->c;
    a.nelt = h->nelt;
    a.xcd_map = h->xcd_map;
    a.reserved = 0;
    a.stamp = h->stamp;
    a.tile_off = h->tile_off;
    a.tile_adj = h->tile_adj;
    a.always = h->always;
    a.gstep = h->gstep;
    a.skip = h->skip;
    a.p = h->kp;
    step_launcher L = pick_launcher(h->prm.iflux, h->fric);
    for (int k = 0; k < nsteps; k++) {
        a.step = k;
        a.st_in = h->st[(h->cur + k) & 1];
        a.st_out = h->st[(h->cur + k + 1) & 1];
        L(h, a, h->nblocks);
        hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1024), 0, h->stream, h->c, h->nblocks, k,
                           h->kp.pad_dt, h->kp.ts, h->kp.tol);
    }
    HIPCHK(hipGetLastError());
    return CF_OK;
}

extern "C" int cf_step_finish(cf_handle *h, double *dt_hist, int *steps_done) {
    if (!h) return CF_ERR_ARG;
    if (steps_done) *steps_done = 0;
    if (h->pending < 0) return CF_ERR_ARG;
    const int nsteps = h->pending;
    h->pending = -1;
    if (nsteps == 0) {
        if (dt_hist) dt_hist[0] = h->dt;
        return CF_OK;
    }
    int hdr[4];
    HIPCHK(hipMemcpyAsync(hdr, h->c.hdr, sizeof(hdr), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const int done = hdr[1];
    std::vector<double> dts((size_t)done + 1), tcs(1);
    std::vector<int> ran((size_t)done + 1, 0);
    HIPCHK(hipMemcpyAsync(dts.data(), h->c.dt, sizeof(double) * ((size_t)done + 1), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(tcs.data(), h->c.tc + done, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (done > 0) HIPCHK(hipMemcpyAsync(ran.data(), h->c.ran, sizeof(int) * (size_t)done, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->last_tiles_run = 0;
    for (int k = 0; k < done; k++) h->last_tiles_run += ran[k];
    h->last_tiles_total = (long long)done * h->nblocks;
    if (dt_hist) memcpy(dt_hist, dts.data(), sizeof(double) * ((size_t)done + 1));
    h->cur = (h->cur + done) & 1;
    h->gstep += (unsigned)done;
    h->dt = dts[done];
    h->tc = tcs[0];
    if (steps_done) *steps_done = done;
    if (hdr[0] < 0) {
        h->status = -1;
        return CF_ERR_DT;
    }
    return CF_OK;
}

extern "C" int cf_step(cf_handle *h, int nsteps, double *dt_hist, int *steps_done) {
    if (steps_done) *steps_done = 0;
    int rc = cf_step_async(h, nsteps);
    if (rc) return rc;
    return cf_step_finish(h, dt_hist, steps_done);
}

extern "C" int cf_set_inflow(cf_handle *h, const double *debitglob, int n) {
    if (!h || !debitglob || n < 0 || n > CF_MAX_IO) return CF_ERR_ARG;
    for (int i = 0; i < n; i++) h->prm.debitglob[i] = debitglob[i];
    fill_kparams(h);
    return CF_OK;
}

extern "C" int cf_set_hydrograph(cf_handle *h, int inlet, int npts, const double *t, const double *q) {
    if (!h || inlet < 1 || inlet > CF_MAX_IO || npts < 0 || npts > CF_MAX_HYDRO || (npts > 0 && (!t || !q)))
        return CF_ERR_ARG;
    for (int i = 0; i + 1 < npts; i++)
        if (!(t[i + 1] > t[i])) return CF_ERR_ARG;
    h->prm.hydro_n[inlet - 1] = npts;
    for (int i = 0; i < npts; i++) {
        h->prm.hydro_t[(inlet - 1) * CF_MAX_HYDRO + i] = t[i];
        h->prm.hydro_q[(inlet - 1) * CF_MAX_HYDRO + i] = q[i];
    }
    fill_kparams(h);
    return upload_hydro(h);
}

extern "C" int cf_set_manning(cf_handle *h, const double *n) {
    if (!h || !n || h->pending >= 0) return CF_ERR_ARG;
    const size_t N = (size_t)h->nelt;
    std::vector<double> m(N);
    for (size_t i = 0; i < N; i++) m[i] = n[h->hperm[i]];
    HIPCHK(hipMemcpyAsync(h->manning, m.data(), N * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

extern "C" int cf_get_geometry(cf_handle *h, double *surf, double *zm, double *cotemin_arr, double *cotemin) {
    if (!h || h->pending >= 0) return CF_ERR_ARG;
    const size_t N = (size_t)h->nelt;
    std::vector<double> s(N), z(N), cot(N);
    HIPCHK(hipMemcpyAsync(s.data(), h->geo + G_SURF * N, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(z.data(), h->zm, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(cot.data(), h->geo + G_COT * N, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double m = 1000000.0;
    for (size_t i = 0; i < N; i++) {
        const size_t o = (size_t)h->hperm[i];
        if (surf) surf[o] = s[i];
        if (zm) zm[o] = z[i];
        if (cotemin_arr) cotemin_arr[o] = cot[i];
        m = std::min(m, cot[i]);
    }
    if (cotemin) *cotemin = m;
    return CF_OK;
}

extern "C" int cf_last_run_stats(cf_handle *h, long long *tiles_run, long long *tiles_total, int *tile_cells) {
    if (!h) return CF_ERR_ARG;
    if (tiles_run) *tiles_run = h->last_tiles_run;
    if (tiles_total) *tiles_total = h->last_tiles_total;
    if (tile_cells) *tile_cells = kBlock;
    return CF_OK;
}

extern "C" void *cf_get_stream(cf_handle *h) { return h ? (void *)h->stream : nullptr; }

extern "C" int cf_set_stream(cf_handle *h, void *stream) {
    if (!h) return CF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
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
