// One cell's time step (flux over its three faces, bed-slope source,
// Manning friction, update, wet/dry clamp) and its reverse.
//
// The forward body is templated on the scalar type (double: the step;
// Dual: the tangent-linear model) and on a loader that supplies states,
// neighbour codes and geometry, so the same operations serve the gfx950
// kernels and a host build of them.  cell_reverse() is the hand-written
// adjoint of cell_update() with the wet/dry switches, wave-speed cases,
// limiter regions and dt frozen as taken in the forward.
//
// Reference: flux_riemann.cuf:4-429 (compute_flux), source_terms.cuf:4-213
// (compute_source_terms), finite_volumes_method.cuf:260-305
// (update_sol_with_flux_and_source).
#pragma once
#include "../../include/cuteflow_hip.h"
#include "cf_device.h"

#pragma clang fp contract(off)

namespace cf {

// geometry arrays (structure of arrays, stride = padded cell count)
enum { G_N1 = 0, G_N2 = 3, G_LN = 6, G_SURF = 9, G_COT = 10, G_XC = 11, G_YC = 12, G_COUNT = 13 };

// Per-instance constants (the reference's `constant` memory,
// global_data_device.cuf, set by data_transfer.cuf:12-96).
struct KParams {
    double gp, tolisec, tol, cfl, ts, constant_dt, override_manning;
    double pad_dt;  // min(1e6, 1e5 if nelt % 1024 != 0): the reduction's padding
    double tol_reg_perm;  // steady-state tolerance (<= 0: never steady)
    int ilimiteur, inlet, outlet, is_dt_constant, is_override_manning;
    int hydro_off;  // offset of this instance's hydrograph tables in the hydro array
    int ad_flags;   // derivative runs: bit 0 = WAF limiter values frozen
    double debitglob[CF_MAX_IO];
    double h_sortie[CF_MAX_IO];
    double long_entree[CF_MAX_IO];
    int hydro_n[CF_MAX_IO];
};

// Neighbour record of a cell: x, y, z = neighbour across face 0..2
// (>= 0: cell index; < 0: -(class | id << 2), class 1 inlet, 2 outlet,
// 3 wall); w bits 2k..2k+1 = face of the neighbour that shares face k's
// edge, bits 6 + 2k + s = second-ring slot s of face k is an adjoint target.
struct I4 {
    int x, y, z, w;
};
struct D4 {
    double x, y, z, w;
};
template <class T>
struct St {
    T h, hu, hv;
    double zm;
};

CF_HD inline int nb_at(const I4 &v, int j) { return j == 0 ? v.x : (j == 1 ? v.y : v.z); }
CF_HD inline double d4_at(const D4 &v, int j) { return j == 0 ? v.x : (j == 1 ? v.y : v.z); }
CF_HD inline int tjv_of(const I4 &v, int j) { return (v.w >> (2 * j)) & 3; }
constexpr int kR2Shift = 6;

// Geometry of one triangle from its vertices (maillage.cuf:132-159,
// :1254-1331, flux_riemann.cuf:50-74): centroid, 1.8f x inradius (cotemin),
// Heron area, bed level zm (vertex z weighted by the vertex-centroid
// distances), and per face k (nodes k -> k+1) the outward normal and length.
CF_HD inline void cell_geometry(const double vx[3], const double vy[3], const double vz[3], double g[G_COUNT],
                                double &zm) {
    const double x1 = vx[0], y1 = vy[0], x2 = vx[1], y2 = vy[1], x3 = vx[2], y3 = vy[2];
    g[G_XC] = ((x1 + x2) + x3) / 3.0;
    g[G_YC] = ((y1 + y2) + y3) / 3.0;
    {
        const double ca = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
        const double cb = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3));
        const double cc = sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1));
        const double pp = ((ca + cb) + cc) / 2.0;
        const double aire = sqrt(((pp * (pp - ca)) * (pp - cb)) * (pp - cc));
        const double cot = (0.5 * aire) / pp;
        g[G_COT] = kCote * cot;
    }
    {
        const double aa = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
        const double bb = sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
        const double cc = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3));
        const double s = ((aa + bb) + cc) / 2.0;
        g[G_SURF] = sqrt(((s * (s - aa)) * (s - bb)) * (s - cc));
        const double xg = ((x1 + x2) + x3) / 3.0, yg = ((y1 + y2) + y3) / 3.0;
        const double d1 = sqrt((x1 - xg) * (x1 - xg) + (y1 - yg) * (y1 - yg));
        const double d2 = sqrt((x2 - xg) * (x2 - xg) + (y2 - yg) * (y2 - yg));
        const double d3 = sqrt((x3 - xg) * (x3 - xg) + (y3 - yg) * (y3 - yg));
        zm = ((d1 * vz[0] + d2 * vz[1]) + d3 * vz[2]) / ((d1 + d2) + d3);
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
        g[G_N1 + j] = n1;
        g[G_N2 + j] = n2;
        g[G_LN + j] = l2;
    }
}

// Inflow discharge of inlet id at time tc: the hydrograph when one is set,
// else debitglob (flux_riemann.cuf:125-139 reads debitglob).
CF_HD inline double inflow_q(const KParams &p, const double *hydro, int id, double tc) {
    const int n = p.hydro_n[id];
    if (n <= 0) return p.debitglob[id];
    const double *t = hydro + p.hydro_off + id * CF_MAX_HYDRO;
    const double *q = hydro + p.hydro_off + CF_MAX_IO * CF_MAX_HYDRO + id * CF_MAX_HYDRO;
    if (tc <= t[0]) return q[0];
    for (int i = 0; i + 1 < n; i++)
        if (tc < t[i + 1]) return q[i] + (q[i + 1] - q[i]) * ((tc - t[i]) / (t[i + 1] - t[i]));
    return q[n - 1];
}

// Flux of a face whose two reconstructed depths are both below eps: the
// Riemann solvers return a zero flux there; same operations as their
// rotation back to the global frame.
template <class T>
CF_HD inline void zero_flux(double n1, double n2, T flux[3]) {
    const T z = 0.0;
    flux[0] = z;
    flux[1] = n1 * z - n2 * z;
    flux[2] = n2 * z + n1 * z;
}

// phys_inflow: below this wetted inlet area (sum of length x depth) the
// inlet gives no discharge (flux_riemann.cuf:128, single-precision literal)
constexpr double kSurfEntreeMin = (double)1e-8f;

// Normal inflow discharge per unit length of an inlet face for 'inflow'
// (flux_riemann.cuf:121: the inlet's discharge over its length) and
// 'phys_inflow' (:128-132: shared in proportion to the face's depth hautl;
// surf_in = the inlet areas of this instance's current state).
template <class T>
CF_HD inline T inlet_normal_q(const KParams &p, const double *hydro, const double *surf_in, int id, double tc,
                              const T &hautl) {
    if (p.inlet == 2) return T(-(inflow_q(p, hydro, id - 1, tc) / p.long_entree[id - 1]));
    const double s = surf_in ? surf_in[id - 1] : 0.0;
    if (s > kSurfEntreeMin) return -((inflow_q(p, hydro, id - 1, tc) * hautl) / s);
    return T(0.0);
}

// New state of cell i after one step of length dt starting at time tc.
// Inflow and outflow discharges of cell i at the start of a step (state u0,
// time tc): the mass flux x length of its inflow faces (debit_entree_arr,
// flux_riemann.cuf:125-126 and 135-136, inlets 'inflow' and 'phys_inflow';
// surf_in: the instance's inlet areas) and of its fixed-level outflow
// faces (debit_sortie_arr, :147-161), each summed over the faces in face
// order.  The same expressions as the boundary fluxes of cell_update.
template <class L>
CF_HD inline void cell_discharge(const L &ld, const KParams &p, const double *hydro, int i, double tc, double &qo,
                                 double &qi, const double *surf_in = nullptr) {
    const double g = p.gp, tls = p.tolisec;
    const St<double> u0 = ld.state(i);
    const I4 nb = ld.nbr(i);
    double qof[3] = {0.0, 0.0, 0.0}, qif[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 3; j++) {
        const int code = nb_at(nb, j);
        if (code >= 0) continue;
        const int cls = (-code) & 3, id = (-code) >> 2;
        const double n1 = ld.geo(G_N1 + j, i), n2 = ld.geo(G_N2 + j, i), ln = ld.geo(G_LN + j, i);
        if (cls == 1 && (p.inlet == 2 || p.inlet == 3)) {
            double dn;
            if (p.inlet == 2) {
                dn = inlet_normal_q<double>(p, hydro, surf_in, id, tc, 0.0);
            } else {
                double hautl, vitul, vitvl;
                rot(tls, n1, n2, u0.h, u0.hu, u0.hv, hautl, vitul, vitvl);
                dn = inlet_normal_q<double>(p, hydro, surf_in, id, tc, hautl);
            }
            qif[j] = (-dn) * ln;
        } else if (cls == 2 && p.outlet == 1) {
            double hautl, vitul, vitvl;
            rot(tls, n1, n2, u0.h, u0.hu, u0.hv, hautl, vitul, vitvl);
            double f0 = 0.0;
            if (vitul >= 0.0 && vitul / sqrt(g * hautl) < 1.0) f0 = dmax(0.0, p.h_sortie[id - 1] - u0.zm) * vitul;
            qof[j] = f0 * ln;
        }
    }
    qo = (qof[0] + qof[1]) + qof[2];
    qi = (qif[0] + qif[1]) + qif[2];
}

// out = (h, hu, hv) after the wet/dry clamp; xout (may be null) = the
// residual before the implicit friction matrix: res = afm1 * x
// (finite_volumes_method.cuf:279-294).
template <class T, int IFLUX, int FRIC, class L>
CF_HD inline void cell_update(const L &ld, const KParams &p, const double *hydro, int i, double dt, double tc,
                              T out[3], T *xout, const double *surf_in = nullptr) {
    const double g = p.gp, tls = p.tolisec;
    const St<T> u0 = ld.state(i);
    const I4 nb = ld.nbr(i);
    const double zml = u0.zm;
    // first ring: the three neighbour states, gathered once
    St<T> unb[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int cc = nb_at(nb, j);
        unb[j] = cc >= 0 ? ld.state(cc) : u0;
    }
    // A dry cell with dry neighbours across its interior faces and no inlet
    // face stays dry, whatever the scheme: every face is dry/dry (zero flux)
    // or a boundary whose flux vanishes with the depth (wall, fixed-level
    // outlet, which takes its wall branch at zero depth, transmissive), the
    // bed source and friction are zero, and the clamp writes (+0, +0, +0):
    // the same bits as the full update, without its face loop.  The
    // residual is returned as zeros (a clamped cell's reverse does not
    // read it).
    {
        bool still = u0.h == 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int cc = nb_at(nb, j);
            still = still && (cc >= 0 ? unb[j].h == 0.0 : ((-cc) & 3) != 1);
        }
        if (still) {
            out[0] = 0.0;
            out[1] = 0.0;
            out[2] = 0.0;
            if (xout) { xout[0] = 0.0; xout[1] = 0.0; xout[2] = 0.0; }
            return;
        }
    }
    T fl[3][3];
    double fn1[3], fn2[3], fln[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const double n1 = ld.geo(G_N1 + j, i);
        const double n2 = ld.geo(G_N2 + j, i);
        const double ln = ld.geo(G_LN + j, i);
        fn1[j] = n1; fn2[j] = n2; fln[j] = ln;
        const int code = nb_at(nb, j);
        T hautl, vitul, vitvl;
        rot(tls, n1, n2, u0.h, u0.hu, u0.hv, hautl, vitul, vitvl);
        T flux[3] = {0.0, 0.0, 0.0};
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
                } else if (p.inlet == 2 || p.inlet == 3) {
                    const T dn = inlet_normal_q<T>(p, hydro, surf_in, id, tc, hautl);
                    flux[0] = dn;
                    flux[1] = n1 * ((dn * dn) / (hautl + tls) + (g * (hautl * hautl)) / 2.0);
                    flux[2] = n2 * ((dn * dn) / (hautl + tls) + (g * (hautl * hautl)) / 2.0);
                }
            } else if (cls == 3) {  // wall (:141-145)
                flux[1] = ((n1 * g) * (hautl * hautl)) / 2.0;
                flux[2] = ((n2 * g) * (hautl * hautl)) / 2.0;
            } else if (cls == 2) {
                if (p.outlet == 1) {  // fixed level (:147-163)
                    if (vitul >= 0.0 && vitul / vsqrt(g * hautl) < 1.0) {
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
            const St<T> ur4 = (kv == i) ? u0 : unb[j];
            const double zmr = ur4.zm;
            const double zij = dmax(zml, zmr);
            T hijet = dmax(T(0.0), (hautl + zml) - zij);
            T hautr, vitur, vitvr;
            rot(tls, n1, n2, ur4.h, ur4.hu, ur4.hv, hautr, vitur, vitvr);
            if (IFLUX == 1 || IFLUX == 2) {  // :197-228: raw right depth
                T hjiet = dmax(T(0.0), (ur4.h + zmr) - zij);
                if (hijet < tls) { hijet = 0.0; vitul = 0.0; vitvl = 0.0; }
                if (hjiet < tls) { hjiet = 0.0; vitur = 0.0; vitvr = 0.0; }
                if (IFLUX == 1) hllc_flux(g, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr, flux);
                else riadh_flux(g, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr, flux);
            } else {
                // reconstructed depths of flux_riemann.cuf:371-394 (the
                // values solver_Riemann_waf / _muscl receive)
                T hjiet = dmax(T(0.0), (hautr + zmr) - zij);
                if (hijet < tls) { hijet = 0.0; vitul = 0.0; vitvl = 0.0; }
                if (hjiet < tls) { hjiet = 0.0; vitur = 0.0; vitvr = 0.0; }
                // dry/dry: no upwind data needed.  WAF: the solver's own
                // test; MUSCL: both depths zero reconstruct and evolve to zero
                const bool dry = IFLUX == 3 ? (hijet < kEpsWaf && hjiet < kEpsWaf)
                                            : (hijet <= 0.0 && hjiet <= 0.0);
                if (dry) {
                    zero_flux(n1, n2, flux);
                } else {
                    // read here, per wet face: read ahead of the face loop,
                    // the centroid distances took registers across the whole
                    // loop and one of them was spilled, the spill waiting for
                    // its load before the geometry loads could start.  The
                    // opaque index keeps the compiler from sharing (and then
                    // spilling) one address across the three faces.
                    const D4 dfi = ld.dfc(cf_opaque(i));
                    const double d_mrl = d4_at(dfi, j);
                    const int ja = (j + 1) % 3, jb = (j + 2) % 3;
                    const bool own1 = nb_at(nb, ja) < 0, own2 = nb_at(nb, jb) < 0;
                    const St<T> s1 = own1 ? u0 : unb[ja];
                    const St<T> s2 = own2 ? u0 : unb[jb];
                    T h1, u1, w1, h2, u2, w2;
                    rot(tls, n1, n2, s1.h, s1.hu, s1.hv, h1, u1, w1);
                    rot(tls, n1, n2, s2.h, s2.hu, s2.hv, h2, u2, w2);
                    const int tjv = tjv_of(nb, j);
                    const I4 nbk = (kv == i) ? nb : ld.nbr(kv);
                    const int ka = (tjv + 1) % 3, kb = (tjv + 2) % 3;
                    const int e1 = nb_at(nbk, ka), e2 = nb_at(nbk, kb);
                    const int kw1 = e1 >= 0 ? e1 : kv, kw2 = e2 >= 0 ? e2 : kv;
                    const St<T> g1 = ld.state(kw1), g2 = ld.state(kw2);
                    T g1h, g1u, g1w, g2h, g2u, g2w;
                    rot(tls, n1, n2, g1.h, g1.hu, g1.hv, g1h, g1u, g1w);
                    rot(tls, n1, n2, g2.h, g2.hu, g2.hv, g2h, g2u, g2w);
                    T hl_upw, vl_upw, hr_upw, vr_upw, ul_upw, ur_upw;  // u_upw: MUSCL only
                    double zml_upw, zmr_upw, d_mkvol, d_mkvoisin;
                    if ((own1 && own2) || (kw1 == kv && kw2 == kv)) {
                        hl_upw = hautl; ul_upw = vitul; vl_upw = vitvl;
                        hr_upw = hautr; ur_upw = vitur; vr_upw = vitvr;
                        d_mkvol = d_mrl; d_mkvoisin = d_mrl;
                        zml_upw = zml; zmr_upw = zmr;
                    } else {
                        if (own1) {
                            hl_upw = h2; ul_upw = u2; vl_upw = w2;
                            d_mkvol = d4_at(dfi, jb);
                            zml_upw = s2.zm;
                        } else if (own2) {
                            hl_upw = h1; ul_upw = u1; vl_upw = w1;
                            d_mkvol = d4_at(dfi, ja);
                            zml_upw = s1.zm;
                        } else {
                            const double d1 = d4_at(dfi, ja), d2 = d4_at(dfi, jb);
                            hl_upw = (d1 * h1 + d2 * h2) / (d1 + d2);
                            ul_upw = (d1 * u1 + d2 * u2) / (d1 + d2);
                            vl_upw = (d1 * w1 + d2 * w2) / (d1 + d2);
                            zml_upw = (d1 * s1.zm + d2 * s2.zm) / (d1 + d2);
                            d_mkvol = (d1 + d2) / 2.0;
                        }
                        const D4 dfk = (kv == i) ? dfi : ld.dfc(kv);
                        if (kw1 == kv) {
                            hr_upw = g2h; ur_upw = g2u; vr_upw = g2w;
                            d_mkvoisin = d4_at(dfk, kb);
                            zmr_upw = g2.zm;
                        } else if (kw2 == kv) {
                            hr_upw = g1h; ur_upw = g1u; vr_upw = g1w;
                            d_mkvoisin = d4_at(dfk, ka);
                            zmr_upw = g1.zm;
                        } else {
                            const double d1 = d4_at(dfk, ka), d2 = d4_at(dfk, kb);
                            hr_upw = (d1 * g1h + d2 * g2h) / (d1 + d2);
                            ur_upw = (d1 * g1u + d2 * g2u) / (d1 + d2);
                            vr_upw = (d1 * g1w + d2 * g2w) / (d1 + d2);
                            zmr_upw = (d1 * g1.zm + d2 * g2.zm) / (d1 + d2);
                            d_mkvoisin = (d1 + d2) / 2.0;
                        }
                    }
                    // upwind hydrostatic reconstruction (flux_riemann.cuf:373,376,384-400)
                    hl_upw = dmax(T(0.0), (hl_upw + zml_upw) - zij);
                    hr_upw = dmax(T(0.0), (hr_upw + zmr_upw) - zij);
                    if (hl_upw < tls) { hl_upw = 0.0; ul_upw = 0.0; vl_upw = 0.0; }
                    if (hr_upw < tls) { hr_upw = 0.0; ur_upw = 0.0; vr_upw = 0.0; }
                    if (IFLUX == 3)
                        waf_flux(g, p.ilimiteur, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr,
                                 hl_upw, vl_upw, hr_upw, vr_upw, d_mrl, d_mkvol, d_mkvoisin, dt, flux,
                                 (p.ad_flags & 1) != 0);
                    else
                        muscl_flux(g, p.ilimiteur, tls, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr,
                                   hl_upw, ul_upw, vl_upw, hr_upw, ur_upw, vr_upw, d_mrl, d_mkvol, d_mkvoisin,
                                   dt, flux);
                }
            }
        }
        fl[j][0] = flux[0] * ln;
        fl[j][1] = flux[1] * ln;
        fl[j][2] = flux[2] * ln;
    }
    // face sum (flux_riemann.cuf:411-426)
    const T resc0 = (fl[0][0] + fl[1][0]) + fl[2][0];
    const T resc1 = (fl[0][1] + fl[1][1]) + fl[2][1];
    const T resc2 = (fl[0][2] + fl[1][2]) + fl[2][2];

    // bed slope source over interior faces (source_terms.cuf:50-100)
    const T q1 = u0.h, q2 = u0.hu, q3 = u0.hv;
    T sb1 = 0.0, sb2 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (nb_at(nb, j) >= 0) {
            const double zij = dmax(zml, unb[j].zm);
            const T hijet = dmax(T(0.0), (q1 + zml) - zij);
            sb1 = sb1 + (((g / 2.0) * (hijet * hijet - q1 * q1)) * fln[j]) * fn1[j];
            sb2 = sb2 + (((g / 2.0) * (hijet * hijet - q1 * q1)) * fln[j]) * fn2[j];
        }
    }
    const double surf = ld.geo(G_SURF, i);
    T r0, r1, r2;
    if (FRIC == 0) {
        // update without friction (finite_volumes_method.cuf:294)
        r0 = (0.0 - resc0) / surf;
        r1 = (sb1 - resc1) / surf;
        r2 = (sb2 - resc2) / surf;
        if (xout) { xout[0] = r0; xout[1] = r1; xout[2] = r2; }
    } else {
        // friction (source_terms.cuf:102-209)
        const double tol = p.tol;
        const T mann = p.is_override_manning ? T(p.override_manning) : ld.manning(i);
        T sf1 = 0.0, sf2 = 0.0, vitu = 0.0, vitv = 0.0, normvit = 0.0;
        if (q1 > tls) {
            vitu = q2 / q1;
            vitv = q3 / q1;
            normvit = vsqrt(vitu * vitu + vitv * vitv);
            if (normvit < tol) normvit = tol;
            const T p13 = cf_pow(q1, -1.0 / 3.0);
            sf1 = -(((((g * mann) * mann) * vitu) * normvit) * surf * p13);
            sf2 = -(((((g * mann) * mann) * vitv) * normvit) * surf * p13);
        }
        const T x0 = ((0.0 + 0.0) - resc0) / surf;
        const T x1 = ((sb1 + sf1) - resc1) / surf;
        const T x2 = ((sb2 + sf2) - resc2) / surf;
        if (xout) { xout[0] = x0; xout[1] = x1; xout[2] = x2; }
        T a0 = 1.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 1.0, a5 = 0.0, a6 = 0.0, a7 = 0.0, a8 = 1.0;
        if (FRIC == 2 && normvit > tol && q1 > tls) {
            const T dtgn2h = ((((dt / 2.0) * g) * mann) * mann) * cf_pow(q1, -4.0 / 3.0);
            const T f21 = -((((7.0 / 3.0) * dtgn2h) * vitu) * normvit);
            const T f22 = 1.0 + (dtgn2h * (2.0 * (vitu * vitu) + vitv * vitv)) / normvit;
            const T f23 = ((dtgn2h * vitu) * vitv) / normvit;
            const T f31 = -((((7.0 / 3.0) * dtgn2h) * vitv) * normvit);
            const T f32 = f23;
            const T f33 = 1.0 + (dtgn2h * (vitu * vitu + 2.0 * (vitv * vitv))) / normvit;
            const T det = f22 * f33 - f23 * f32;
            a3 = -(f21 * f33 - f23 * f31) / det;
            a4 = f33 / det;
            a5 = -f23 / det;
            a6 = (f21 * f32 - f22 * f31) / det;
            a7 = -f32 / det;
            a8 = f22 / det;
        } else if (FRIC == 3 && normvit > tol && q1 > tls) {
            const T beta = -(((((dt * g) * mann) * mann) * normvit) * cf_pow(q1, -4.0 / 3.0));
            const T det = (1.0 - beta) * (1.0 - beta);
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
    T h = u0.h + dt * r0;
    T hu = u0.hu + dt * r1;
    T hv = u0.hv + dt * r2;
    if (h <= tls) { h = 0.0; hu = 0.0; hv = 0.0; }
    out[0] = h;
    out[1] = hu;
    out[2] = hv;
}

// ---------------------------------------------------------------------------
// Reverse of cell_update for one cell.
//
// Inputs: the step's inputs (through the loader), x = the residual the
// forward computed (xout of cell_update), lam = cotangent of the new state
// (h, hu, hv) of cell i.  Outputs (overwritten):
//   own    cotangent of cell i's own old state;
//   r1[k]  cotangent of the old state of neighbour nb_k(i) (k = 0..2);
//   r2[k][s] cotangent of the old state of the second-ring cell in slot s
//          of face k (WAF upwind cells across the face);
//   gn     cotangent of Manning n of cell i.
// The state cotangents of all cells are the sums of these contributions
// over the cells whose update read them.

// physical flux (h u, n1 e - n2 t, n2 e + n1 t), e = h u^2 + g h^2/2,
// t = h u v: cotangents of (h, u, v)
CF_HD inline void phys_flux_rev(double g, double n1, double n2, double H, double U, double V, const double bf[3],
                                double &bH, double &bU, double &bV) {
    const double be = n1 * bf[1] + n2 * bf[2];
    const double bt = -n2 * bf[1] + n1 * bf[2];
    bH += bf[0] * U + be * (U * U + g * H) + bt * (U * V);
    bU += bf[0] * H + be * (2.0 * H * U) + bt * (H * V);
    bV += bt * (H * U);
}

template <class V3>
CF_HD inline void add3(V3 &a, double x, double y, double z) {
    a[0] += x;
    a[1] += y;
    a[2] += z;
}

template <int IFLUX, int FRIC, class L>
CF_HD inline void cell_reverse(const L &ld, const KParams &p, const double *hydro, int i, double dt, double tc,
                               const double x[3], const double lam[3], double own[3], double r1[3][3],
                               double r2[3][2][3], double &gn) {
    const double g = p.gp, tls = p.tolisec;
    for (int k = 0; k < 3; k++) {
        own[k] = 0.0;
        for (int c = 0; c < 3; c++) r1[k][c] = 0.0;
        for (int s = 0; s < 2; s++)
            for (int c = 0; c < 3; c++) r2[k][s][c] = 0.0;
    }
    gn = 0.0;
    const St<double> u0 = ld.state(i);
    const double q1 = u0.h, q2 = u0.hu, q3 = u0.hv;
    const double surf = ld.geo(G_SURF, i);
    const double x0 = x[0], x1 = x[1], x2 = x[2];
    // the neighbour codes and states, read before the friction part so that
    // their loads overlap it
    const I4 nb = ld.nbr(i);
    St<double> unb[3];
    for (int j = 0; j < 3; j++) {
        const int cc = nb_at(nb, j);
        unb[j] = cc >= 0 ? ld.state(cc) : u0;
    }

    // friction values and the implicit matrix (as in cell_update)
    double mann = 0.0, sf1 = 0.0, sf2 = 0.0, vitu = 0.0, vitv = 0.0, normvit = 0.0, p13 = 0.0;
    bool floored = false;
    double a3 = 0.0, a4 = 1.0, a5 = 0.0, a6 = 0.0, a7 = 0.0, a8 = 1.0;
    bool mat = false;
    double dtgn2h = 0.0, f21 = 0.0, f22 = 0.0, f23 = 0.0, f31 = 0.0, f32 = 0.0, f33 = 0.0, det = 0.0, beta = 0.0;
    double p43 = 0.0;
    if (FRIC != 0) {
        const double tol = p.tol;
        mann = p.is_override_manning ? p.override_manning : ld.manning(i);
        if (q1 > tls) {
            vitu = q2 / q1;
            vitv = q3 / q1;
            normvit = sqrt(vitu * vitu + vitv * vitv);
            if (normvit < tol) {
                normvit = tol;
                floored = true;
            }
            p13 = cf_pow(q1, -1.0 / 3.0);
            sf1 = -(((((g * mann) * mann) * vitu) * normvit) * surf * p13);
            sf2 = -(((((g * mann) * mann) * vitv) * normvit) * surf * p13);
        }
        if (FRIC == 2 && normvit > tol && q1 > tls) {
            mat = true;
            // q1^(-4/3) = q1^(-1/3) / q1: enters only linear coefficients here
            p43 = p13 / q1;
            dtgn2h = ((((dt / 2.0) * g) * mann) * mann) * p43;
            f21 = -((((7.0 / 3.0) * dtgn2h) * vitu) * normvit);
            f22 = 1.0 + (dtgn2h * (2.0 * (vitu * vitu) + vitv * vitv)) / normvit;
            f23 = ((dtgn2h * vitu) * vitv) / normvit;
            f31 = -((((7.0 / 3.0) * dtgn2h) * vitv) * normvit);
            f32 = f23;
            f33 = 1.0 + (dtgn2h * (vitu * vitu + 2.0 * (vitv * vitv))) / normvit;
            det = f22 * f33 - f23 * f32;
            a3 = -(f21 * f33 - f23 * f31) / det;
            a4 = f33 / det;
            a5 = -f23 / det;
            a6 = (f21 * f32 - f22 * f31) / det;
            a7 = -f32 / det;
            a8 = f22 / det;
        } else if (FRIC == 3 && normvit > tol && q1 > tls) {
            mat = true;
            p43 = p13 / q1;
            beta = -(((((dt * g) * mann) * mann) * normvit) * p43);
            det = (1.0 - beta) * (1.0 - beta);
            a4 = (1.0 - beta) / det;
            a8 = (1.0 - beta) / det;
        }
    }
    (void)sf1;
    (void)sf2;
    // the clamp: r0 = x0 for every friction mode
    const double r0 = (FRIC == 0 || FRIC == 1) ? x0 : (1.0 * x0 + 0.0 * x1) + 0.0 * x2;
    const double hnew = u0.h + dt * r0;
    if (hnew <= tls) return;  // clamped: the new state is the constant 0

    own[0] = lam[0];
    own[1] = lam[1];
    own[2] = lam[2];
    const double br0 = dt * lam[0], br1 = dt * lam[1], br2 = dt * lam[2];
    double bx0 = br0, bx1 = br1, bx2 = br2;
    double bq1 = 0.0, bq2 = 0.0, bq3 = 0.0, bm = 0.0;
    double bvitu = 0.0, bvitv = 0.0, bnv = 0.0;
    if (FRIC == 2 || FRIC == 3) {
        bx0 = br0 + a3 * br1 + a6 * br2;
        bx1 = a4 * br1 + a7 * br2;
        bx2 = a5 * br1 + a8 * br2;
    }
    if (FRIC == 2 && mat) {
        const double ba3 = br1 * x0, ba4 = br1 * x1, ba5 = br1 * x2;
        const double ba6 = br2 * x0, ba7 = br2 * x1, ba8 = br2 * x2;
        const double idet = 1.0 / det;
        const double bdet = -(ba3 * a3 + ba4 * a4 + ba5 * a5 + ba6 * a6 + ba7 * a7 + ba8 * a8) * idet;
        const double bN3 = ba3 * idet, bN4 = ba4 * idet, bN5 = ba5 * idet;
        const double bN6 = ba6 * idet, bN7 = ba7 * idet, bN8 = ba8 * idet;
        double bf21 = 0.0, bf22 = 0.0, bf23 = 0.0, bf31 = 0.0, bf32 = 0.0, bf33 = 0.0;
        // a3 = -(f21 f33 - f23 f31)/det, a4 = f33/det, a5 = -f23/det,
        // a6 = (f21 f32 - f22 f31)/det, a7 = -f32/det, a8 = f22/det
        bf21 += -bN3 * f33 + bN6 * f32;
        bf33 += -bN3 * f21 + bN4 + bdet * f22;
        bf23 += bN3 * f31 - bN5 - bdet * f32;
        bf31 += bN3 * f23 - bN6 * f22;
        bf32 += bN6 * f21 - bN7 - bdet * f23;
        bf22 += -bN6 * f31 + bN8 + bdet * f33;
        bf23 += bf32;  // f32 = f23
        const double D = dtgn2h, u = vitu, v = vitv, w = normvit, iw = 1.0 / normvit;
        const double c73 = 7.0 / 3.0;
        const double e22 = 2.0 * (u * u) + v * v, e33 = u * u + 2.0 * (v * v);
        const double bD = -c73 * (bf21 * u * w + bf31 * v * w) + (bf22 * e22 + bf23 * (u * v) + bf33 * e33) * iw;
        const double Diw = D * iw;
        bvitu += -c73 * bf21 * D * w + Diw * (bf22 * (4.0 * u) + bf23 * v + bf33 * (2.0 * u));
        bvitv += -c73 * bf31 * D * w + Diw * (bf22 * (2.0 * v) + bf23 * u + bf33 * (4.0 * v));
        bnv += -c73 * (bf21 * D * u + bf31 * D * v) - (bf22 * e22 + bf23 * (u * v) + bf33 * e33) * (Diw * iw);
        // D = (dt/2 g m m) q1^(-4/3)
        bm += bD * (((dt / 2.0) * g) * (2.0 * mann)) * p43;
        bq1 += bD * D * (-4.0 / 3.0) / q1;
    } else if (FRIC == 3 && mat) {
        // a4 = a8 = 1/(1 - beta)
        const double ba = br1 * x1 + br2 * x2;
        const double bbeta = ba * (a4 * a4);
        // beta = -(dt g m m w q1^(-4/3))
        bm += bbeta * (-(((dt * g) * (2.0 * mann)) * normvit) * p43);
        bnv += bbeta * (-((((dt * g) * mann) * mann) * p43));
        bq1 += bbeta * beta * (-4.0 / 3.0) / q1;
    }
    // x = ((sb + sf) - resc) / surf
    const double isurf = 1.0 / surf;
    const double bres[3] = {-bx0 * isurf, -bx1 * isurf, -bx2 * isurf};
    const double bs1 = bx1 * isurf, bs2 = bx2 * isurf;  // sb and sf
    if (FRIC != 0 && q1 > tls) {
        // sf = -(g m m) u w surf p13
        const double K = (g * mann) * mann;
        bm += -(bs1 * vitu + bs2 * vitv) * ((((g * 2.0) * mann) * normvit) * surf * p13);
        bvitu += -bs1 * (K * normvit * surf * p13);
        bvitv += -bs2 * (K * normvit * surf * p13);
        bnv += -(bs1 * vitu + bs2 * vitv) * (K * surf * p13);
        const double bp13 = -(bs1 * vitu + bs2 * vitv) * (K * normvit * surf);
        const double iq1 = 1.0 / q1;
        bq1 += bp13 * (-1.0 / 3.0) * (p13 * iq1);
        if (!floored) {
            const double inv = 1.0 / normvit;
            bvitu += bnv * (vitu * inv);
            bvitv += bnv * (vitv * inv);
        }
        bq2 += bvitu * iq1;
        bq3 += bvitv * iq1;
        bq1 -= (bvitu * vitu + bvitv * vitv) * iq1;
    }
    gn = bm;

    const double zml = u0.zm;
    // bed slope source: sb += (g/2 (hij^2 - q1^2) ln) n over interior faces
    for (int j = 0; j < 3; j++) {
        if (nb_at(nb, j) >= 0) {
            const double zij = dmax(zml, unb[j].zm);
            const double raw = (q1 + zml) - zij;
            const double hij = dmax(0.0, raw);
            const double dh = (!(0.0 > raw) ? hij : 0.0) - q1;
            bq1 += (bs1 * ld.geo(G_N1 + j, i) + bs2 * ld.geo(G_N2 + j, i)) * (g * dh) * ld.geo(G_LN + j, i);
        }
    }
    own[0] += bq1;
    own[1] += bq2;
    own[2] += bq3;

    // faces: resc = sum of flux * ln
    D4 dfi = {0.0, 0.0, 0.0, 0.0};
    if (IFLUX >= 3) dfi = ld.dfc(i);
    // unrolled: with a runtime face index the neighbour states and the
    // contribution rows live in private (scratch) memory
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const double n1 = ld.geo(G_N1 + j, i);
        const double n2 = ld.geo(G_N2 + j, i);
        const double ln = ld.geo(G_LN + j, i);
        const double bf[3] = {bres[0] * ln, bres[1] * ln, bres[2] * ln};
        const int code = nb_at(nb, j);
        double hautl, vitul, vitvl;
        rot(tls, n1, n2, q1, q2, q3, hautl, vitul, vitvl);
        double bhl = 0.0, bul = 0.0, bvl = 0.0;  // cotangents of (hautl, vitul, vitvl)
        int kv = -1;
        if (code >= 0) {
            kv = code;
        } else {
            const int cls = (-code) & 3, id = (-code) >> 2;
            if (cls == 1) {
                if (p.inlet == 1) {
                    phys_flux_rev(g, n1, n2, hautl, vitul, vitvl, bf, bhl, bul, bvl);
                } else if (p.inlet == 2) {
                    const double dn = -(inflow_q(p, hydro, id - 1, tc) / p.long_entree[id - 1]);
                    const double be = n1 * bf[1] + n2 * bf[2];
                    bhl += be * (-(dn * dn) / ((hautl + tls) * (hautl + tls)) + g * hautl);
                }
            } else if (cls == 3) {
                bhl += (n1 * bf[1] + n2 * bf[2]) * (g * hautl);
            } else if (cls == 2) {
                if (p.outlet == 1) {
                    if (vitul >= 0.0 && vitul / sqrt(g * hautl) < 1.0) {
                        const double hout = dmax(0.0, p.h_sortie[id - 1] - zml);
                        double bdummy = 0.0;
                        phys_flux_rev(g, n1, n2, hout, vitul, vitvl, bf, bdummy, bul, bvl);
                    } else {
                        bhl += (n1 * bf[1] + n2 * bf[2]) * (g * hautl);
                    }
                } else {
                    kv = i;
                }
            }
        }
        // cotangent of the right state (cell kv) in the face frame: raw h, hautr, un, ut
        double bhr_raw = 0.0, bhr = 0.0, bur = 0.0, bvr = 0.0;
        if (kv >= 0) {
            const St<double> ur4 = (kv == i) ? u0 : unb[j];
            const double zmr = ur4.zm;
            const double zij = dmax(zml, zmr);
            const double rawl = (hautl + zml) - zij;
            const double hijet = dmax(0.0, rawl);
            double hautr, vitur, vitvr;
            rot(tls, n1, n2, ur4.h, ur4.hu, ur4.hv, hautr, vitur, vitvr);
            const double hjiet = dmax(0.0, (IFLUX <= 2 ? (ur4.h + zmr) - zij : (hautr + zmr) - zij));
            const bool zl = hijet < tls, zr = hjiet < tls;
            const double fhl = zl ? 0.0 : hijet, ful = zl ? 0.0 : vitul, fvl = zl ? 0.0 : vitvl;
            const double fhr = zr ? 0.0 : hjiet, fur = zr ? 0.0 : vitur, fvr = zr ? 0.0 : vitvr;
            const bool actl = !zl && !(0.0 > rawl);  // dmax(0, raw) returns raw on a tie
            const double rawr2 = (IFLUX <= 2) ? (ur4.h + zmr) - zij : (hautr + zmr) - zij;
            const bool actr = !zr && !(0.0 > rawr2);
            if (IFLUX == 1 || IFLUX == 2) {
                double gg[6] = {0, 0, 0, 0, 0, 0};
                if (IFLUX == 1) hllc_flux_rev(g, n1, n2, fhl, ful, fvl, fhr, fur, fvr, bf, gg);
                else riadh_flux_rev(g, n1, n2, fhl, ful, fvl, fhr, fur, fvr, bf, gg);
                if (!zl) {
                    if (actl) bhl += gg[0];
                    bul += gg[1];
                    bvl += gg[2];
                }
                if (!zr) {
                    if (actr) bhr_raw += gg[3];
                    bur += gg[4];
                    bvr += gg[5];
                }
            } else if (IFLUX == 3 ? !(fhl < kEpsWaf && fhr < kEpsWaf) : !(fhl <= 0.0 && fhr <= 0.0)) {
                // WAF (iflux 3) or MUSCL (iflux 4) on the two-ring stencil
                const double d_mrl = d4_at(dfi, j);
                const int ja = (j + 1) % 3, jb = (j + 2) % 3;
                const bool own1 = nb_at(nb, ja) < 0, own2 = nb_at(nb, jb) < 0;
                const St<double> s1 = own1 ? u0 : unb[ja];
                const St<double> s2 = own2 ? u0 : unb[jb];
                double h1, u1, w1, h2, u2, w2;
                rot(tls, n1, n2, s1.h, s1.hu, s1.hv, h1, u1, w1);
                rot(tls, n1, n2, s2.h, s2.hu, s2.hv, h2, u2, w2);
                const int tjv = tjv_of(nb, j);
                const I4 nbk = (kv == i) ? nb : ld.nbr(kv);
                const int ka = (tjv + 1) % 3, kb = (tjv + 2) % 3;
                const int e1 = nb_at(nbk, ka), e2 = nb_at(nbk, kb);
                const int kw1 = e1 >= 0 ? e1 : kv, kw2 = e2 >= 0 ? e2 : kv;
                const St<double> g1 = ld.state(kw1), g2 = ld.state(kw2);
                double g1h, g1u, g1w, g2h, g2u, g2w;
                rot(tls, n1, n2, g1.h, g1.hu, g1.hv, g1h, g1u, g1w);
                rot(tls, n1, n2, g2.h, g2.hu, g2.hv, g2h, g2u, g2w);
                const bool special = (own1 && own2) || (kw1 == kv && kw2 == kv);
                double hl0, ul0, vl0, hr0, ur0, vr0, zml_upw, zmr_upw, d_mkvol, d_mkvoisin;
                double wl1 = 0.0, wl2 = 0.0, wr1 = 0.0, wr2 = 0.0;  // weights of s1, s2, g1, g2
                if (special) {
                    hl0 = hautl; ul0 = ful; vl0 = fvl;
                    hr0 = hautr; ur0 = fur; vr0 = fvr;
                    d_mkvol = d_mrl; d_mkvoisin = d_mrl;
                    zml_upw = zml; zmr_upw = zmr;
                } else {
                    if (own1) {
                        hl0 = h2; ul0 = u2; vl0 = w2; d_mkvol = d4_at(dfi, jb); zml_upw = s2.zm; wl2 = 1.0;
                    } else if (own2) {
                        hl0 = h1; ul0 = u1; vl0 = w1; d_mkvol = d4_at(dfi, ja); zml_upw = s1.zm; wl1 = 1.0;
                    } else {
                        const double d1 = d4_at(dfi, ja), d2 = d4_at(dfi, jb);
                        hl0 = (d1 * h1 + d2 * h2) / (d1 + d2);
                        ul0 = (d1 * u1 + d2 * u2) / (d1 + d2);
                        vl0 = (d1 * w1 + d2 * w2) / (d1 + d2);
                        zml_upw = (d1 * s1.zm + d2 * s2.zm) / (d1 + d2);
                        d_mkvol = (d1 + d2) / 2.0;
                        wl1 = d1 / (d1 + d2);
                        wl2 = d2 / (d1 + d2);
                    }
                    const D4 dfk = (kv == i) ? dfi : ld.dfc(kv);
                    if (kw1 == kv) {
                        hr0 = g2h; ur0 = g2u; vr0 = g2w; d_mkvoisin = d4_at(dfk, kb); zmr_upw = g2.zm; wr2 = 1.0;
                    } else if (kw2 == kv) {
                        hr0 = g1h; ur0 = g1u; vr0 = g1w; d_mkvoisin = d4_at(dfk, ka); zmr_upw = g1.zm; wr1 = 1.0;
                    } else {
                        const double d1 = d4_at(dfk, ka), d2 = d4_at(dfk, kb);
                        hr0 = (d1 * g1h + d2 * g2h) / (d1 + d2);
                        ur0 = (d1 * g1u + d2 * g2u) / (d1 + d2);
                        vr0 = (d1 * g1w + d2 * g2w) / (d1 + d2);
                        zmr_upw = (d1 * g1.zm + d2 * g2.zm) / (d1 + d2);
                        d_mkvoisin = (d1 + d2) / 2.0;
                        wr1 = d1 / (d1 + d2);
                        wr2 = d2 / (d1 + d2);
                    }
                }
                const double rawlu = (hl0 + zml_upw) - zij, rawru = (hr0 + zmr_upw) - zij;
                double hlu = dmax(0.0, rawlu), hru = dmax(0.0, rawru);
                const bool zlu = hlu < tls, zru = hru < tls;
                const double fhlu = zlu ? 0.0 : hlu, fulu = zlu ? 0.0 : ul0, fvlu = zlu ? 0.0 : vl0;
                const double fhru = zru ? 0.0 : hru, furu = zru ? 0.0 : ur0, fvru = zru ? 0.0 : vr0;
                // gg: face states (0..5), then the upwind (h, u, v) left and right
                double gg[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                if (IFLUX == 3) {
                    double gw[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                    waf_flux_rev(g, p.ilimiteur, n1, n2, fhl, ful, fvl, fhr, fur, fvr, fhlu, fvlu, fhru, fvru, d_mrl,
                                 d_mkvol, d_mkvoisin, dt, bf, gw, (p.ad_flags & 1) != 0);
                    for (int q = 0; q < 6; q++) gg[q] = gw[q];
                    gg[6] = gw[6];
                    gg[8] = gw[7];
                    gg[9] = gw[8];
                    gg[11] = gw[9];
                } else {
                    muscl_flux_rev(g, p.ilimiteur, tls, n1, n2, fhl, ful, fvl, fhr, fur, fvr, fhlu, fulu, fvlu, fhru,
                                   furu, fvru, d_mrl, d_mkvol, d_mkvoisin, dt, bf, gg);
                }
                // face states
                double bul_post = 0.0, bvl_post = 0.0, bur_post = 0.0, bvr_post = 0.0;
                if (!zl) {
                    if (actl) bhl += gg[0];
                    bul += gg[1];
                    bvl_post += gg[2];
                }
                if (!zr) {
                    if (actr) bhr += gg[3];
                    bur += gg[4];
                    bvr_post += gg[5];
                }
                // upwind states
                const double bhl0 = (!zlu && !(0.0 > rawlu)) ? gg[6] : 0.0, bul0 = zlu ? 0.0 : gg[7],
                             bvl0 = zlu ? 0.0 : gg[8];
                const double bhr0 = (!zru && !(0.0 > rawru)) ? gg[9] : 0.0, bur0 = zru ? 0.0 : gg[10],
                             bvr0 = zru ? 0.0 : gg[11];
                if (special) {
                    bhl += bhl0;
                    bul_post += bul0;
                    bvl_post += bvl0;
                    bhr += bhr0;
                    bur_post += bur0;
                    bvr_post += bvr0;
                } else {
                    // left upwind: s1 = neighbour ja, s2 = neighbour jb (never own here)
                    if (wl1 != 0.0) {
                        double bh = 0.0, bhu = 0.0, bhv = 0.0;
                        rot_rev(tls, n1, n2, s1.h, s1.hu, s1.hv, bhl0 * wl1, bul0 * wl1, bvl0 * wl1, bh, bhu, bhv);
                        add3(r1[ja], bh, bhu, bhv);
                    }
                    if (wl2 != 0.0) {
                        double bh = 0.0, bhu = 0.0, bhv = 0.0;
                        rot_rev(tls, n1, n2, s2.h, s2.hu, s2.hv, bhl0 * wl2, bul0 * wl2, bvl0 * wl2, bh, bhu, bhv);
                        add3(r1[jb], bh, bhu, bhv);
                    }
                    // right upwind: second-ring slots 0 (kw1) and 1 (kw2)
                    if (wr1 != 0.0) {
                        double bh = 0.0, bhu = 0.0, bhv = 0.0;
                        rot_rev(tls, n1, n2, g1.h, g1.hu, g1.hv, bhr0 * wr1, bur0 * wr1, bvr0 * wr1, bh, bhu, bhv);
                        add3(r2[j][0], bh, bhu, bhv);
                    }
                    if (wr2 != 0.0) {
                        double bh = 0.0, bhu = 0.0, bhv = 0.0;
                        rot_rev(tls, n1, n2, g2.h, g2.hu, g2.hv, bhr0 * wr2, bur0 * wr2, bvr0 * wr2, bh, bhu, bhv);
                        add3(r2[j][1], bh, bhu, bhv);
                    }
                }
                if (!zl) {
                    bul += bul_post;
                    bvl += bvl_post;
                }
                if (!zr) {
                    bur += bur_post;
                    bvr += bvr_post;
                }
            }
            // the right state's cotangent goes to cell kv
            double bh = bhr_raw, bhu = 0.0, bhv = 0.0;
            rot_rev(tls, n1, n2, ur4.h, ur4.hu, ur4.hv, bhr, bur, bvr, bh, bhu, bhv);
            if (kv == i) add3(own, bh, bhu, bhv);
            else add3(r1[j], bh, bhu, bhv);
        }
        // the left state (cell i)
        double bh = 0.0, bhu = 0.0, bhv = 0.0;
        rot_rev(tls, n1, n2, q1, q2, q3, bhl, bul, bvl, bh, bhu, bhv);
        add3(own, bh, bhu, bhv);
    }
}

}  // namespace cf
