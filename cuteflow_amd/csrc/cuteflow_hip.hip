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

__device__ double inflow_q(const KParams &p, const double *hydro, int id, double tc) {
    const int n = p.hydro_n[id];
    if (n <= 0) return p.debitglob[id];
    const double *t = hydro + p.hydro_off + id * CF_MAX_HYDRO;
    const double *q = hydro + p.hydro_off + CF_MAX_IO * CF_MAX_HYDRO + id * CF_MAX_HYDRO;
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
__global__ __launch_bounds__(kTile) CF_STEP_OCCUPANCY void step_kernel(StepArgs a) {
    __shared__ double s_min[kWaves];
    const int tile = blockIdx.x;
    const int b = a.tile_inst[tile];
    const InstCtrl c = a.ctrl[b];
    if (c.status != 0) {  // this instance has stopped (ts reached) or failed
        if (threadIdx.x == 0) a.partial[tile] = __builtin_inf();
        return;
    }
    // A tile whose cells and neighbour tiles' cells were all dry after the
    // last two steps would recompute (0, 0, 0) for every cell (all faces
    // dry/dry, no source, clamp), and both state buffers already hold that:
    // it is skipped.  Reading a stamp raised concurrently only makes the
    // tile run, which is always correct.
    if (a.skip && !a.always[tile] && a.stamp[tile] + 1u < c.gstep) {
        if (threadIdx.x == 0) a.partial[tile] = __builtin_inf();  // no dt contribution
        return;
    }
    const double4 *__restrict__ st_in = a.in_x ? a.in_x : a.st[(c.par + a.step) & 1];
    double4 *__restrict__ st_out = a.out_x ? a.out_x : a.st[(c.par + a.step + 1) & 1];
    const int nelt = a.ncell;  // geometry stride
    const int i = tile * kTile + threadIdx.x;
    const double dt = c.dt;
    const double tc = c.tc;
    const KParams &p = a.kp[b];
    const double g = p.gp, tls = p.tolisec;
    double dti = 1000000.0;
    int isnan_local = 0, wet_local = 0;

    if (i < a.inst_end[b]) {
        const double4 u0 = st_in[i];
        const int4 nb = a.nbr[i];
        const double zml = u0.w;
        // first ring: the three neighbour states, gathered once
        double4 unb[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int cc = nb_at(nb, j);
            unb[j] = cc >= 0 ? st_in[cc] : u0;
        }
        double4 dfi = make_double4(0.0, 0.0, 0.0, 0.0);
        if (IFLUX == 3) dfi = a.dfc[i];
        double fl[3][3], fn1[3], fn2[3], fln[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double n1 = a.geo[(G_N1 + j) * (size_t)nelt + i];
            const double n2 = a.geo[(G_N2 + j) * (size_t)nelt + i];
            const double ln = a.geo[(G_LN + j) * (size_t)nelt + i];
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
                        const double4 g1 = st_in[kw1], g2 = st_in[kw2];
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
        const double surf = a.geo[G_SURF * (size_t)nelt + i];
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
        st_out[i] = make_double4(h, hu, hv, zml);
        isnan_local = (h != h) || (hu != hu) || (hv != hv);
        wet_local = (h != 0.0) || isnan_local;
        // cflcond on the new state (cfl_condition.cuf:20-36)
        if (h > tls) {
            const double uu = hu / h, vv = hv / h;
            const double nv = sqrt(uu * uu + vv * vv);
            const double vmax = dmax(0.01, nv + sqrt(g * fabs(h)));
            dti = dmin((p.cfl * a.geo[G_COT * (size_t)nelt + i]) / vmax, 1000000.0);
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
        a.partial[tile] = anynan ? __builtin_nan("") : bm;
    }
    // a wet tile keeps itself and its neighbour tiles running for two steps
    if (a.skip && __syncthreads_or(wet_local)) {
        const int o = a.tile_adj_off[tile], n = a.tile_adj_off[tile + 1] - o;
        for (int k = threadIdx.x; k <= n; k += kTile)
            a.stamp[k == 0 ? tile : a.tile_adj[o + k - 1]] = c.gstep + 1u;
    }
}

// Reduces per-tile minima (NaN marks a NaN state, +inf a skipped tile)
// starting from `pad`; every thread gets the result, the NaN flag and the
// number of skipped tiles.
__device__ double block_min_1024(const double *__restrict__ part, int n, double pad, int &nan, int &nskip) {
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

// End of a step, one workgroup per instance: dt of the new state (the
// reference's reduction, cfl_condition.cuf:68-104), tc += dt, the dt range
// check (:129-132) and the loop condition (finite_volumes_method.cuf:123).
__global__ __launch_bounds__(1024) void finalize_kernel(InstCtrl *ctrl, const KParams *kp, const double *partial,
                                                         const int *inst_tile, double *dt_hist, double *tc_hist,
                                                         int *ran, int hist_stride, int step) {
    const int b = blockIdx.x;
    InstCtrl *c = ctrl + b;
    if (c->status != 0) return;
    const int t0 = inst_tile[b], nt = inst_tile[b + 1] - t0;
    int nan, nskip;
    const double dtn = block_min_1024(partial + t0, nt, kp[b].pad_dt, nan, nskip);
    if (threadIdx.x == 0) {
        const double tcn = c->tc + c->dt;
        c->dt = dtn;
        c->tc = tcn;
        c->steps = step + 1;
        c->gstep += 1u;
        dt_hist[(size_t)b * hist_stride + step + 1] = dtn;
        tc_hist[(size_t)b * hist_stride + step + 1] = tcn;
        ran[(size_t)b * hist_stride + step] = nt - nskip;
        if (nan || !(dtn <= 9999.0 && dtn >= kDtLow)) c->status = -1;
        else if (!(((tcn - kp[b].ts) + 2.0 * kp[b].tol) < kp[b].tol)) c->status = 1;
    }
}

// Start of a run: the buffer parity moves past the previous run's steps,
// the loop condition is evaluated (a failed instance stays failed).
__global__ void run_init_kernel(InstCtrl *ctrl, const KParams *kp, double *dt_hist, double *tc_hist,
                                int hist_stride, int ninst) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ninst) return;
    InstCtrl *c = ctrl + b;
    c->par = (c->par + c->steps) & 1;
    c->steps = 0;
    if (c->status >= 0) c->status = (((c->tc - kp[b].ts) + 2.0 * kp[b].tol) < kp[b].tol) ? 0 : 1;
    dt_hist[(size_t)b * hist_stride] = c->dt;
    tc_hist[(size_t)b * hist_stride] = c->tc;
}

// cflcond for the current state of one instance, one minimum per tile
// (cfl_condition.cuf:4-39); like the reference, the initial dt does not
// look for NaN states.
__global__ __launch_bounds__(kTile) void cfl_kernel(const double4 *__restrict__ st, const double *__restrict__ cot,
                                                     int cell0, int cell_end, const KParams *kp, int b,
                                                     double *__restrict__ part) {
    __shared__ double s_min[kWaves];
    const int i = cell0 + blockIdx.x * kTile + threadIdx.x;
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
        for (int w = 1; w < kWaves; w++) bm = dmin(bm, s_min[w]);
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
    const double x1 = vx[0], y1 = vy[0], x2 = vx[1], y2 = vy[1], x3 = vx[2], y3 = vy[2];
    const size_t S = (size_t)stride;
    geo[G_XC * S + i] = ((x1 + x2) + x3) / 3.0;
    geo[G_YC * S + i] = ((y1 + y2) + y3) / 3.0;
    {
        const double ca = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
        const double cb = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3));
        const double cc = sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1));
        const double pp = ((ca + cb) + cc) / 2.0;
        const double aire = sqrt(((pp * (pp - ca)) * (pp - cb)) * (pp - cc));
        const double cot = (0.5 * aire) / pp;
        geo[G_COT * S + i] = kCote * cot;
    }
    {
        const double aa = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
        const double bb = sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
        const double cc = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3));
        const double s = ((aa + bb) + cc) / 2.0;
        geo[G_SURF * S + i] = sqrt(((s * (s - aa)) * (s - bb)) * (s - cc));
        const double xg = ((x1 + x2) + x3) / 3.0, yg = ((y1 + y2) + y3) / 3.0;
        const double d1 = sqrt((x1 - xg) * (x1 - xg) + (y1 - yg) * (y1 - yg));
        const double d2 = sqrt((x2 - xg) * (x2 - xg) + (y2 - yg) * (y2 - yg));
        const double d3 = sqrt((x3 - xg) * (x3 - xg) + (y3 - yg) * (y3 - yg));
        zm_out[i] = ((d1 * vz[0] + d2 * vz[1]) + d3 * vz[2]) / ((d1 + d2) + d3);
    }
    for (int j = 0; j < 3; j++) {
        const int c1 = j, c2 = (j + 1) % 3, c3 = (j + 2) % 3;
        const double ax = vx[c1], ay = vy[c1], bx = vx[c2], by = vy[c2], cx = vx[c3], cy = vy[c3];
        const double l2 = sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay));
        double n1 = (by - ay) / l2, n2 = (ax - bx) / l2;
        const double orient = n1 * ((ax + bx) / 2.0 - ((ax + bx) + cx) / 3.0) +
                              n2 * ((ay + by) / 2.0 - ((ay + by) + cy) / 3.0);
        if (orient < 0.0) {
            n1 = (ay - by) / l2;
            n2 = (bx - ax) / l2;
        }
        geo[(G_N1 + j) * S + i] = n1;
        geo[(G_N2 + j) * S + i] = n2;
        geo[(G_LN + j) * S + i] = l2;
    }
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
        const int k = nb_at(nb, j);
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

__global__ void fill_u32_kernel(unsigned *p, int n, unsigned v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void pow_kernel(const double *__restrict__ x, double y, int n, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cf_pow(x[i], y);
}

}  // namespace

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "cuteflow_hip: %s failed: %s\n", #x, hipGetErrorString(e_));  \
            return CF_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

static int grid_of(int n, int b) { return (n + b - 1) / b; }

template <typename T>
static void dfree(T *&ptr) {
    if (ptr) (void)hipFree((void *)ptr);
    ptr = nullptr;
}

template <typename T>
static bool dalloc(T *&p, size_t count) {
    return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)) == hipSuccess;
}

template <typename T>
static bool h2d(T *dst, const T *src, size_t count, hipStream_t s) {
    return count == 0 || hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s) == hipSuccess;
}

static KParams make_kparams(const cf_params &p, int nelt, int hydro_off) {
    KParams k{};
    k.gp = p.gp; k.tolisec = p.tolisec; k.tol = p.tol; k.cfl = p.cfl; k.ts = p.ts;
    k.constant_dt = p.constant_dt; k.override_manning = p.override_manning;
    k.pad_dt = (nelt % 1024 != 0) ? 100000.0 : 1000000.0;
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
    hipLaunchKernelGGL(cfl_kernel, dim3(nt), dim3(kTile), 0, h->stream, h->st[cur], h->geo + G_COT * (size_t)h->ncell,
                       c0, c0 + h->nelt[b], h->d_kp, b, part);
    hipLaunchKernelGGL(reset_inst_kernel, dim3(1), dim3(1024), 0, h->stream, part, nt, h->d_kp, b, h->ctrl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->hctrl.data() + b, h->ctrl + b, sizeof(InstCtrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CF_OK;
}

static int ensure_hist(cf_handle *h, int nsteps) {
    if (nsteps <= h->cap) return CF_OK;
    const int cap = std::max(nsteps, 2 * h->cap);
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h->dt_hist); dfree(h->tc_hist); dfree(h->ran);
    drop_graphs(h);
    const size_t n = (size_t)h->ninst * (size_t)(cap + 1);
    if (!dalloc(h->dt_hist, n) || !dalloc(h->tc_hist, n) || !dalloc(h->ran, n)) return CF_ERR_NOMEM;
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
    drop_graphs(h);
    dfree(h->st[0]); dfree(h->st[1]); dfree(h->nbr); dfree(h->geo); dfree(h->dfc);
    dfree(h->manning); dfree(h->hydro); dfree(h->zm); dfree(h->stage); dfree(h->perm);
    dfree(h->d_kp); dfree(h->ctrl); dfree(h->partial); dfree(h->stamp);
    dfree(h->tile_adj_off); dfree(h->tile_adj); dfree(h->always);
    dfree(h->tile_inst); dfree(h->inst_end); dfree(h->inst_tile);
    dfree(h->dt_hist); dfree(h->tc_hist); dfree(h->ran);
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
    std::vector<cfprep::Prep> prep(ninst);
    for (int b = 0; b < ninst; b++) {
        int st = cfprep::prepare(meshes + b, params + b, kTile, (params[b].options & CF_OPT_KEEP_ORDER) != 0, prep[b]);
        if (st) return st;
    }
    cf_handle *h = new (std::nothrow) cf_handle();
    if (!h) return CF_ERR_NOMEM;
    h->ninst = ninst;
    h->iflux = params[0].iflux;
    h->fric = params[0].friction == 0 ? 0 : 1 + params[0].fricimplic;
    h->nelt.resize(ninst); h->nnt.resize(ninst); h->cell_off.resize(ninst); h->tile_off.resize(ninst + 1);
    h->prm.assign(params, params + ninst);
    h->kp.resize(ninst);
    h->hperm.resize(ninst);
    h->hctrl.assign(ninst, InstCtrl{});
    h->last_run.assign(ninst, 0);
    h->last_total.assign(ninst, 0);
    // skipping dry tiles is exact except with a constant dt (every cell then
    // contributes constant_dt to the minimum)
    h->skip = (params[0].options & CF_OPT_NO_TILE_SKIP) ? 0 : 1;
    for (int b = 0; b < ninst; b++)
        if (params[b].is_dt_constant == 1) h->skip = 0;
    int cells = 0, tiles = 0;
    for (int b = 0; b < ninst; b++) {
        h->nelt[b] = meshes[b].nelt;
        h->nnt[b] = meshes[b].nnt;
        h->cell_off[b] = cells;
        h->tile_off[b] = tiles;
        const int nt = grid_of(meshes[b].nelt, kTile);
        tiles += nt;
        cells += nt * kTile;
        h->max_cells = std::max(h->max_cells, meshes[b].nelt);
        h->kp[b] = make_kparams(params[b], meshes[b].nelt, b * kHydroPerInst);
        h->hperm[b] = prep[b].perm;
    }
    h->tile_off[ninst] = tiles;
    h->ncell = cells;
    h->ntiles = tiles;
    const size_t N = (size_t)cells;

    // global (batched) host tables
    std::vector<int4> nbr(N, make_int4(-3, -3, -3, 0));
    std::vector<double> mann(N, 0.0);
    std::vector<int> perm(N, 0), tile_inst(tiles), inst_end(ninst), inst_tile(ninst + 1), adj_off(tiles + 1, 0), adj;
    std::vector<unsigned char> always(tiles, 0);
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
            !dalloc(h->inst_tile, (size_t)ninst + 1)) {
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
    *out = h;
    return CF_OK;
}

extern "C" int cf_create(const cf_mesh *mesh, const cf_params *params, cf_handle **out) {
    return cf_batch_create(1, mesh, params, out);
}

extern "C" void cf_destroy(cf_handle *h) {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
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
static void launch_step(const StepArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((step_kernel<IFLUX, FRIC>), dim3(grid), dim3(kTile), 0, s, a);
}

typedef void (*step_launcher)(const StepArgs &, int, hipStream_t);

static step_launcher pick_launcher(int iflux, int fric) {
    static const step_launcher tab[2][4] = {
        {launch_step<1, 0>, launch_step<1, 1>, launch_step<1, 2>, launch_step<1, 3>},
        {launch_step<3, 0>, launch_step<3, 1>, launch_step<3, 2>, launch_step<3, 3>}};
    return tab[iflux == 3 ? 1 : 0][fric];
}

static StepArgs base_args(cf_handle *h) {
    StepArgs a{};
    a.st[0] = h->st[0];
    a.st[1] = h->st[1];
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
    a.partial = h->partial;
    a.stamp = h->stamp;
    a.tile_adj_off = h->tile_adj_off;
    a.tile_adj = h->tile_adj;
    a.always = h->always;
    a.tile_inst = h->tile_inst;
    a.inst_end = h->inst_end;
    a.inst_tile = h->inst_tile;
    a.in_x = nullptr;
    a.out_x = nullptr;
    a.ncell = h->ncell;
    a.hist_stride = h->cap + 1;
    a.skip = h->skip;
    return a;
}

// The launches of a run of nsteps steps, on stream s.
static void enqueue_run(cf_handle *h, int nsteps, hipStream_t s) {
    hipLaunchKernelGGL(run_init_kernel, dim3(grid_of(h->ninst, 256)), dim3(256), 0, s, h->ctrl, h->d_kp, h->dt_hist,
                       h->tc_hist, h->cap + 1, h->ninst);
    StepArgs a = base_args(h);
    step_launcher L = pick_launcher(h->iflux, h->fric);
    for (int k = 0; k < nsteps; k++) {
        a.step = k;
        L(a, h->ntiles, s);
        hipLaunchKernelGGL(finalize_kernel, dim3(h->ninst), dim3(1024), 0, s, h->ctrl, h->d_kp, h->partial,
                           h->inst_tile, h->dt_hist, h->tc_hist, h->ran, h->cap + 1, k);
    }
}

// A run's launches are captured once per step count and replayed: every
// argument is fixed (per-run state lives in the device control block).
static int run_graph(cf_handle *h, int nsteps) {
    for (auto &g : h->graphs)
        if (g.nsteps == nsteps) {
            HIPCHK(hipGraphLaunch(g.exec, h->stream));
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
    h->graphs.push_back({nsteps, {}, exec});
    HIPCHK(hipGraphLaunch(exec, h->stream));
    return CF_OK;
}

extern "C" int cf_step_async(cf_handle *h, int nsteps) {
    if (!h || nsteps < 0) return CF_ERR_ARG;
    if (h->pending >= 0) return CF_ERR_ARG;  // previous run not finished
    bool any = false;
    for (int b = 0; b < h->ninst; b++) any = any || h->hctrl[b].status >= 0;
    if (!any) return CF_ERR_DT;
    int rc = ensure_hist(h, std::max(nsteps, 1));
    if (rc) return rc;
    h->pending = nsteps;
    if (nsteps == 0) return CF_OK;
    // the null stream cannot be captured
    if (h->use_graphs && h->stream && nsteps <= 4096) return run_graph(h, nsteps);
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
    std::vector<InstCtrl> c(B);
    HIPCHK(hipMemcpyAsync(c.data(), h->ctrl, B * sizeof(InstCtrl), hipMemcpyDeviceToHost, h->stream));
    std::vector<double> dts((size_t)B * S);
    std::vector<int> ran((size_t)B * S);
    HIPCHK(hipMemcpyAsync(dts.data(), h->dt_hist, dts.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(ran.data(), h->ran, ran.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    int rc = CF_OK;
    for (int b = 0; b < B; b++) {
        const int done = c[b].steps;
        if (dt_hist) memcpy(dt_hist + (size_t)b * (nsteps + 1), dts.data() + (size_t)b * S, sizeof(double) * ((size_t)done + 1));
        if (steps_done) steps_done[b] = done;
        if (status) status[b] = c[b].status < 0 ? CF_ERR_DT : CF_OK;
        if (c[b].status < 0) rc = CF_ERR_DT;
        long long run = 0;
        for (int k = 0; k < done; k++) run += ran[(size_t)b * S + k];
        h->last_run[b] = run;
        h->last_total[b] = (long long)done * (h->tile_off[b + 1] - h->tile_off[b]);
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

extern "C" int cf_last_run_stats(cf_handle *h, long long *tiles_run, long long *tiles_total, int *tile_cells) {
    if (!h) return CF_ERR_ARG;
    long long r = 0, t = 0;
    for (int b = 0; b < h->ninst; b++) {
        r += h->last_run[b];
        t += h->last_total[b];
    }
    if (tiles_run) *tiles_run = r;
    if (tiles_total) *tiles_total = t;
    if (tile_cells) *tile_cells = kTile;
    return CF_OK;
}

extern "C" int cf_batch_set_inflow(cf_handle *h, int b, const double *debitglob, int n) {
    if (bad_inst(h, b) || !debitglob || n < 0 || n > CF_MAX_IO) return CF_ERR_ARG;
    for (int i = 0; i < n; i++) h->prm[b].debitglob[i] = debitglob[i];
    h->kp[b] = make_kparams(h->prm[b], h->nelt[b], b * kHydroPerInst);
    return upload_params(h, b);
}

extern "C" int cf_set_inflow(cf_handle *h, const double *debitglob, int n) { return cf_batch_set_inflow(h, 0, debitglob, n); }

extern "C" int cf_batch_set_hydrograph(cf_handle *h, int b, int inlet, int npts, const double *t, const double *q) {
    if (bad_inst(h, b) || inlet < 1 || inlet > CF_MAX_IO || npts < 0 || npts > CF_MAX_HYDRO || (npts > 0 && (!t || !q)))
        return CF_ERR_ARG;
    for (int i = 0; i + 1 < npts; i++)
        if (!(t[i + 1] > t[i])) return CF_ERR_ARG;
    h->prm[b].hydro_n[inlet - 1] = npts;
    for (int i = 0; i < npts; i++) {
        h->prm[b].hydro_t[(inlet - 1) * CF_MAX_HYDRO + i] = t[i];
        h->prm[b].hydro_q[(inlet - 1) * CF_MAX_HYDRO + i] = q[i];
    }
    h->kp[b] = make_kparams(h->prm[b], h->nelt[b], b * kHydroPerInst);
    return upload_params(h, b);
}

extern "C" int cf_set_hydrograph(cf_handle *h, int inlet, int npts, const double *t, const double *q) {
    return cf_batch_set_hydrograph(h, 0, inlet, npts, t, q);
}

extern "C" int cf_batch_set_manning(cf_handle *h, int b, const double *n) {
    if (bad_inst(h, b) || !n) return CF_ERR_ARG;
    const size_t N = (size_t)h->nelt[b];
    std::vector<double> m(N);
    for (size_t i = 0; i < N; i++) m[i] = n[h->hperm[b][i]];
    HIPCHK(hipMemcpyAsync(h->manning + h->cell_off[b], m.data(), N * sizeof(double), hipMemcpyHostToDevice, h->stream));
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
