// Numerics of the St-Venant step: project pow, HLLC and WAF Riemann solvers,
// limiter, rotation to the face frame; and their reverse (adjoint) forms.
//
// Every expression keeps the operation order of the reference Fortran
// (flux_riemann.cuf, source_terms.cuf) so that the results are bit-identical
// to a CPU evaluation of the same formulas; translation units are compiled
// with -ffp-contract=off.  The forward functions are templated on the scalar
// type: double for the step, Dual for the tangent-linear model (same value
// bits, plus the directional derivative).  The *_rev functions are the
// hand-written reverse of the forward ones for fixed branches: given the
// cotangents of the outputs they accumulate the cotangents of the inputs.
#pragma once
#include <stdint.h>

#include "cf_dual.h"
#include "cf_hd.h"

#pragma clang fp contract(off)

namespace cf {

// Default-real literals of the reference are single precision, widened.
constexpr double kEpsWaf = (double)1e-16f;   // flux_riemann.cuf:749
constexpr double kEpsHllc = (double)1e-8f;   // flux_riemann.cuf:926
constexpr double kEpsRiadh = (double)1e-6f;  // flux_riemann.cuf:1051
constexpr double kQSteady = (double)1e-8f;   // finite_volumes_method.cuf:379
constexpr double kDtLow = (double)1e-12f;    // cfl_condition.cuf:129
constexpr double kEpsLeer = (double)1e-12f;  // flux_riemann.cuf:453
constexpr double kCote = (double)1.8f;       // maillage.cuf:157

template <class T>
CF_HD inline T dmax(T a, T b) { return a > b ? a : b; }
template <class T>
CF_HD inline T dmin(T a, T b) { return a < b ? a : b; }

CF_HD inline void two_prod(double a, double b, double &p, double &e) {
    const double split = 134217729.0;  // 2^27 + 1 (Veltkamp)
    p = a * b;
    double t = split * a;
    double ah = t - (t - a), al = a - ah;
    t = split * b;
    double bh = t - (t - b), bl = b - bh;
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}

CF_HD inline void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// x^y for x > 0: double-double log (atanh series) and Cody-Waite exp.
// Uses only +,-,*,/, floor and bit operations: same bits as the host build.
CF_HD inline double cf_pow(double x, double y) {
    const double ln2_hi = 0x1.62e42ffp-1;
    const double ln2_lo = -0x1.718432a1b0e26p-35;
    const double inv_ln2 = 0x1.71547652b82fep+0;
    const double sqrt2 = 0x1.6a09e667f3bcdp+0;
    if (!(x > 0.0) || x != x || y != y) return __builtin_nan("");
    if (x == __builtin_inf()) return y > 0 ? __builtin_inf() : (y < 0 ? 0.0 : 1.0);
    uint64_t u = __builtin_bit_cast(uint64_t, x);
    int e = 0;
    if ((u >> 52) == 0) {
        u = __builtin_bit_cast(uint64_t, x * 0x1p54);
        e = -54;
    }
    e += (int)((u >> 52) & 0x7ff) - 1023;
    u = (u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m = __builtin_bit_cast(double, u);
    if (m >= sqrt2) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double d = 2.0 + f;
    double dlo = (2.0 - d) + f;
    double s = f / d;
    double p, plo;
    two_prod(s, d, p, plo);
    double r = ((f - p) - plo) - s * dlo;
    double slo = r / d;
    double s2 = s * s;
    double t = 0x1.47ae147ae147bp-5;
    t = t * s2 + 0x1.642c8590b2164p-5;
    t = t * s2 + 0x1.8618618618618p-5;
    t = t * s2 + 0x1.af286bca1af28p-5;
    t = t * s2 + 0x1.e1e1e1e1e1e1ep-5;
    t = t * s2 + 0x1.1111111111111p-4;
    t = t * s2 + 0x1.3b13b13b13b14p-4;
    t = t * s2 + 0x1.745d1745d1746p-4;
    t = t * s2 + 0x1.c71c71c71c71cp-4;
    t = t * s2 + 0x1.2492492492492p-3;
    t = t * s2 + 0x1.999999999999ap-3;
    t = t * s2 + 0x1.5555555555555p-2;
    t = t * s2;
    double lh, ll, tmp;
    two_sum((double)e * ln2_hi, 2.0 * s, lh, ll);
    ll = ll + ((double)e * ln2_lo + (2.0 * slo + 2.0 * s * t));
    tmp = lh + ll;
    ll = ll - (tmp - lh);
    lh = tmp;
    double th, tl;
    two_prod(y, lh, th, tl);
    tl = tl + y * ll;
    tmp = th + tl;
    tl = tl - (tmp - th);
    th = tmp;
    if (th > 709.0) return __builtin_inf();
    if (th < -745.0) return 0.0;
    int k = (int)floor(th * inv_ln2 + 0.5);
    double rh = th - (double)k * ln2_hi;
    r = (rh - (double)k * ln2_lo) + tl;
    double q = 0x1.6124613a86d09p-33;
    q = q * r + 0x1.1eed8eff8d898p-29;
    q = q * r + 0x1.ae64567f544e4p-26;
    q = q * r + 0x1.27e4fb7789f5cp-22;
    q = q * r + 0x1.71de3a556c734p-19;
    q = q * r + 0x1.a01a01a01a01ap-16;
    q = q * r + 0x1.a01a01a01a01ap-13;
    q = q * r + 0x1.6c16c16c16c17p-10;
    q = q * r + 0x1.1111111111111p-7;
    q = q * r + 0x1.5555555555555p-5;
    q = q * r + 0x1.5555555555555p-3;
    q = q * r + 0.5;
    double ex = 1.0 + (r + (r * r) * q);
    int k1 = k / 2, k2 = k - k1;
    ex = ex * __builtin_bit_cast(double, (uint64_t)(k1 + 1023) << 52);
    ex = ex * __builtin_bit_cast(double, (uint64_t)(k2 + 1023) << 52);
    return ex;
}

// d(x^y) = y x^y / x dx
CF_HD inline Dual cf_pow(Dual x, double y) {
    const double p = cf_pow(x.v, y);
    return Dual(p, x.d * ((y * p) / x.v));
}

// Wave speeds and HLLC flux in the face frame (flux_riemann.cuf:930-1020 for
// HLLC, :753-845 for the first half of WAF).  Returns false for a dry/dry
// face: the reference leaves al, ar, pql, pqr unset there; the flux is zero.
template <class T>
struct HllcT {
    T f1[3], f2[3], fh[3];
    T sl, sr, set;
};

template <class T>
CF_HD inline bool hllc_core(double gp, double eps, T hl, T ul, T vl, T hr, T ur, T vr, HllcT<T> &o) {
    if (hl < eps && hr < eps) return false;
    T al = vsqrt(gp * hl);
    T ar = vsqrt(gp * hr);
    T het = (hl + hr) / 2.0 - (((ur - ul) * (hl + hr)) / (al + ar)) / 4.0;
    T uet = (ul + ur) / 2.0 - ((hr - hl) * (al + ar)) / (hl + hr);
    T pql, pqr, sl, sr, set;
    if (het < hl) pql = 1.0;
    else pql = (hl > eps) ? vsqrt(((0.5 * (het + hl)) * het) / (hl * hl)) : T(0.0);
    if (het < hr) pqr = 1.0;
    else pqr = (hr > eps) ? vsqrt(((0.5 * (het + hr)) * het) / (hr * hr)) : T(0.0);
    if (hl > eps) {
        sl = ul - al * pql;
        if (hr > eps) {
            sr = ur + ar * pqr;
        } else {
            sl = ul - al;
            sr = ul + 2.0 * al;
        }
    } else {
        sl = ur - 2.0 * ar;
        sr = ur + ar;
    }
    T denom = hr * (ur - sr) - hl * (ul - sl);
    if (vabs(denom) < eps) set = uet;
    else set = ((sl * hr) * (ur - sr) - (sr * hl) * (ul - sl)) / denom;
    T q1[3] = {hl, hl * ul, hl * vl};
    T q2[3] = {hr, hr * ur, hr * vr};
    o.f1[0] = hl * ul;
    o.f1[1] = hl * (ul * ul) + ((gp * hl) * hl) / 2.0;
    o.f1[2] = (hl * ul) * vl;
    o.f2[0] = hr * ur;
    o.f2[1] = hr * (ur * ur) + ((gp * hr) * hr) / 2.0;
    o.f2[2] = (hr * ur) * vr;
    if (sl > 0.0) {
        o.fh[0] = o.f1[0]; o.fh[1] = o.f1[1]; o.fh[2] = o.f1[2];
    } else if (sr < 0.0) {
        o.fh[0] = o.f2[0]; o.fh[1] = o.f2[1]; o.fh[2] = o.f2[2];
    } else {
#pragma unroll
        for (int k = 0; k < 2; k++)
            o.fh[k] = ((sr * o.f1[k] - sl * o.f2[k]) + (sl * sr) * (q2[k] - q1[k])) /
                      ((sr - sl) + eps);
        o.fh[2] = (set > 0.0) ? o.fh[0] * vl : o.fh[0] * vr;
    }
    o.sl = sl;
    o.sr = sr;
    o.set = set;
    return true;
}

// solver_Riemann_hllc (flux_riemann.cuf:902-1026): flux in the global frame.
template <class T>
CF_HD inline void hllc_flux(double gp, double n1, double n2, T hl, T ul, T vl, T hr, T ur, T vr, T flux[3]) {
    HllcT<T> o;
    T fh0 = 0.0, fh1 = 0.0, fh2 = 0.0;
    if (hllc_core(gp, kEpsHllc, hl, ul, vl, hr, ur, vr, o)) {
        fh0 = o.fh[0]; fh1 = o.fh[1]; fh2 = o.fh[2];
    }
    flux[0] = fh0;
    flux[1] = n1 * fh1 - n2 * fh2;
    flux[2] = n2 * fh1 + n1 * fh2;
}

// solver_Riemann_riadh (flux_riemann.cuf:1029-1182), iflux 2: the HLLC wave
// speeds with eps = 1e-6f, then the flux of the region holding x/t = 0
// among U_L, U*_L, U*_R, U_R, the star fluxes being F_K + S_K (U*_K - U_K)
// with U*_K = h_K (S_K - u_K)/(S_K - S*) (1, S*, v_K) (weight 0 when
// |S_K - S*| <= eps).  Both sides dry (:1055 leaves al, ar, pql, pqr
// unset): zero flux.
template <class T>
CF_HD inline void riadh_flux(double gp, double n1, double n2, T hl, T ul, T vl, T hr, T ur, T vr, T flux[3]) {
    const double eps = kEpsRiadh;
    if (hl < eps && hr < eps) {
        flux[0] = 0.0;
        flux[1] = 0.0;
        flux[2] = 0.0;
        return;
    }
    T al = vsqrt(gp * hl);
    T ar = vsqrt(gp * hr);
    T het = (hl + hr) / 2.0 - (((ur - ul) * (hl + hr)) / (al + ar)) / 4.0;
    T uet = (ul + ur) / 2.0 - ((hr - hl) * (al + ar)) / (hl + hr);
    T pql, pqr, sl, sr, set;
    if (het < hl) pql = 1.0;
    else pql = (hl > eps) ? vsqrt(((0.5 * (het + hl)) * het) / (hl * hl)) : T(0.0);
    if (het < hr) pqr = 1.0;
    else pqr = (hr > eps) ? vsqrt(((0.5 * (het + hr)) * het) / (hr * hr)) : T(0.0);
    if (hl > eps) {
        sl = ul - al * pql;
        if (hr > eps) {
            sr = ur + ar * pqr;
        } else {
            sl = ul - al;
            sr = ul + 2.0 * al;
        }
    } else {
        sl = ur - 2.0 * ar;
        sr = ur + ar;
    }
    T denom = hr * (ur - sr) - hl * (ul - sl);
    if (vabs(denom) < eps) set = uet;
    else set = ((sl * hr) * (ur - sr) - (sr * hl) * (ul - sl)) / denom;
    T fx[3];
    if (0.0 < sl) {
        fx[0] = hl * ul;
        fx[1] = hl * (ul * ul) + ((gp * hl) * hl) / 2.0;
        fx[2] = (hl * ul) * vl;
    } else if (0.0 < set && 0.0 > sl) {
        const T pond = (vabs(sl - set) > eps) ? (hl * (sl - ul)) / (sl - set) : T(0.0);
        fx[0] = hl * ul + sl * (pond - hl);
        fx[1] = (hl * (ul * ul) + ((gp * hl) * hl) / 2.0) + sl * (pond * set - hl * ul);
        fx[2] = (hl * ul) * vl + sl * (pond * vl - hl * vl);
    } else if (0.0 > set && 0.0 < sr) {
        const T pond = (vabs(sr - set) > eps) ? (hr * (sr - ur)) / (sr - set) : T(0.0);
        fx[0] = hr * ur + sr * (pond - hr);
        fx[1] = (hr * (ur * ur) + ((gp * hr) * hr) / 2.0) + sr * (pond * set - hr * ur);
        fx[2] = (hr * ur) * vr + sr * (pond * vr - hr * vr);
    } else {
        fx[0] = hr * ur;
        fx[1] = hr * (ur * ur) + ((gp * hr) * hr) / 2.0;
        fx[2] = (hr * ur) * vr;
    }
    flux[0] = fx[0];
    flux[1] = n1 * fx[1] - n2 * fx[2];
    flux[2] = n2 * fx[1] + n1 * fx[2];
}

// limiter (flux_riemann.cuf:465-524); unknown ilimiteur -> 1.
template <class T>
CF_HD inline T limiter(int ilim, T r, T c) {
    T ac = vabs(c);
    switch (ilim) {
    case 0: return 1.0;
    case 1:
        if (r <= 0.0) return 1.0;
        else if (r > 0.0 && r < 0.5) return 1.0 - (2.0 * (1.0 - ac)) * r;
        else if (r > 0.5 && r < 1.0) return ac;
        else if (r > 1.0 && r < 2.0) return 1.0 - (1.0 - ac) * r;
        else return 2.0 * ac - 1.0;
    case 2:
        if (r <= 0.0) return 1.0;
        return 1.0 - ((2.0 * (1.0 - ac)) * r) / (1.0 + r);
    case 3:
        if (r <= 0.0) return 1.0;
        return 1.0 - (((1.0 - ac) * r) * (1.0 + r)) / (1.0 + r * r);
    case 4:
        if (r <= 0.0) return 1.0;
        else if (r <= 1.0) return 1.0 - (1.0 - ac) * r;
        else return ac;
    default: return 1.0;
    }
}

// solver_Riemann_waf (flux_riemann.cuf:720-900).
// freeze_lim (derivative runs only): the limiter values are taken as
// constants, like the other switches of the forward; values are unchanged.
template <class T>
CF_HD inline void waf_flux(double gp, int ilim, double n1, double n2, T hl, T ul, T vl, T hr, T ur, T vr,
                           T hl_upw, T vl_upw, T hr_upw, T vr_upw, double d_mrl, double d_mkvol,
                           double d_mkvoisin, double dt, T flux[3], bool freeze_lim = false) {
    const double eps = kEpsWaf;
    HllcT<T> o;
    T fw0 = 0.0, fw1 = 0.0, fw2 = 0.0;
    if (hllc_core(gp, eps, hl, ul, vl, hr, ur, vr, o)) {
        double dtoverdx = dt / d_mrl;
        T cr = o.sr * dtoverdx;
        T cl = o.sl * dtoverdx;
        T cet = o.set * dtoverdx;
        T dlr = (hr - hl) / d_mrl;
        T dll = (hr - hl) / d_mrl;
        T dlet = (vr - vl) / d_mrl;
        T dur = (o.sr > 0.0) ? (hl - hl_upw) / d_mkvol : (hr_upw - hr) / d_mkvoisin;
        T dul = (o.sl > 0.0) ? (hl - hl_upw) / d_mkvol : (hr_upw - hr) / d_mkvoisin;
        T duet = (o.set > 0.0) ? (vl - vl_upw) / d_mkvol : (vr_upw - vr) / d_mkvoisin;
        T rl = dul / (dll + eps);
        T rr = dur / (dlr + eps);
        T ret = duet / (dlet + eps);
        T lim_l = limiter(ilim, rl, cl);
        T lim_r = limiter(ilim, rr, cr);
        T lim_et = limiter(ilim, ret, cet);
        if (freeze_lim) {
            lim_l = T(val(lim_l));
            lim_r = T(val(lim_r));
            lim_et = T(val(lim_et));
        }
        double sgl = copysign(1.0, val(o.sl)), sgr = copysign(1.0, val(o.sr)), sge = copysign(1.0, val(o.set));
        T wl = (1.0 + sgl * lim_l) / 2.0;
        T wlet = (1.0 + sge * lim_et) / 2.0;
        T wlr = (sgr * lim_r - sgl * lim_l) / 2.0;
        T wr = (1.0 - sgr * lim_r) / 2.0;
        T wret = (1.0 - sge * lim_et) / 2.0;
        fw0 = (wl * o.f1[0] + wlr * o.fh[0]) + wr * o.f2[0];
        fw1 = (wl * o.f1[1] + wlr * o.fh[1]) + wr * o.f2[1];
        fw2 = (wlet * vl + wret * vr) * fw0;
    }
    flux[0] = fw0;
    flux[1] = n1 * fw1 - n2 * fw2;
    flux[2] = n2 * fw1 + n1 * fw2;
}

// simple_limiter (flux_riemann.cuf:430-462): the limited slope of the MUSCL
// reconstruction from the one-sided slopes r1 (upwind) and r2 (across the
// face).  ilimiteur 0: none (first order), 1: superbee, 2: van Leer,
// 3: minmod; the reference leaves the slope unset for other values, taken
// as 0 (first order) here.
template <class T>
CF_HD inline T slope_limiter(int ilim, T r1, T r2) {
    switch (ilim) {
    case 1:
        if (r2 > 0.0) return dmax(dmax(T(0.0), dmin(2.0 * r1, r2)), dmin(r1, 2.0 * r2));
        return dmin(dmin(T(0.0), dmax(2.0 * r1, r2)), dmax(r1, 2.0 * r2));
    case 2: return (r1 * r2 + vabs(r1 * r2)) / ((r1 + r2) + kEpsLeer);
    case 3:
        if (r2 > 0.0) return dmax(T(0.0), dmin(r1, r2));
        return dmin(T(0.0), dmax(r1, r2));
    default: return 0.0;
    }
}

// solver_Riemann_muscl (flux_riemann.cuf:526-718), iflux 4: MUSCL-Hancock.
// Each side's cell is reconstructed to its two face values along the face
// normal (limited slope from its upwind value and the other side; zero
// velocities when either reconstructed depth is below tolisec), evolved
// half a step with the difference of its face fluxes, and the evolved
// inner values go to the HLLC solver with eps = 1e-16f.  The evolution of
// the velocity components with momentum-flux differences is the
// reference's.  Both sides dry after the evolution: zero flux.
template <class T>
CF_HD inline void muscl_flux(double gp, int ilim, double tolisec, double n1, double n2, T hl, T ul, T vl, T hr,
                             T ur, T vr, T hl_upw, T ul_upw, T vl_upw, T hr_upw, T ur_upw, T vr_upw,
                             double d_mrl, double d_mkvol, double d_mkvoisin, double dt, T flux[3]) {
    const double dtoverdx = dt / d_mrl;
    // data reconstruction I, left cell (:560-584)
    T hll, hlr, ull, ulr, vll, vlr;
    {
        const T dh = slope_limiter(ilim, (hl - hl_upw) / d_mkvol, (hr - hl) / d_mrl);
        hll = dmax(T(0.0), hl - (0.5 * dh) * d_mrl);
        hlr = dmax(T(0.0), hl + (0.5 * dh) * d_mrl);
        if (hlr < tolisec || hll < tolisec) {
            ull = ulr = vll = vlr = 0.0;
        } else {
            const T du = slope_limiter(ilim, (ul - ul_upw) / d_mkvol, (ur - ul) / d_mrl);
            const T dv = slope_limiter(ilim, (vl - vl_upw) / d_mkvol, (vr - vl) / d_mrl);
            ull = ul - (0.5 * du) * d_mrl;
            vll = vl - (0.5 * dv) * d_mrl;
            ulr = ul + (0.5 * du) * d_mrl;
            vlr = vl + (0.5 * dv) * d_mrl;
        }
    }
    // right cell (:586-610)
    T hrl, hrr, url, urr, vrl, vrr;
    {
        const T dh = slope_limiter(ilim, (hr - hl) / d_mrl, (hr_upw - hr) / d_mkvoisin);
        hrr = dmax(T(0.0), hr + (0.5 * dh) * d_mrl);
        hrl = dmax(T(0.0), hr - (0.5 * dh) * d_mrl);
        if (hrl < tolisec || hrr < tolisec) {
            url = urr = vrl = vrr = 0.0;
        } else {
            const T du = slope_limiter(ilim, (ur - ul) / d_mrl, (ur_upw - ur) / d_mkvoisin);
            const T dv = slope_limiter(ilim, (vr - vl) / d_mrl, (vr_upw - vr) / d_mkvoisin);
            urr = ur + (0.5 * du) * d_mrl;
            vrr = vr + (0.5 * dv) * d_mrl;
            url = ur - (0.5 * du) * d_mrl;
            vrl = vr - (0.5 * dv) * d_mrl;
        }
    }
    // data reconstruction II: half-step evolution (:613-620)
    const double hdt = 0.5 * dtoverdx, hg = 0.5 * gp;
    const T eh = hlr + hdt * (hll * ull - hlr * ulr);
    const T eu = ulr + hdt * ((hll * (ull * ull) + hg * (hll * hll)) - (hlr * (ulr * ulr) + hg * (hlr * hlr)));
    const T ev = vlr + hdt * ((hll * ull) * vll - (hlr * ulr) * vlr);
    const T fh = hrl + hdt * (hrl * url - hrr * urr);
    const T fu = url + hdt * ((hrl * (url * url) + hg * (hrl * hrl)) - (hrr * (urr * urr) + hg * (hrr * hrr)));
    const T fv = vrl + hdt * ((hrl * url) * vrl - (hrr * urr) * vrr);
    HllcT<T> o;
    T fh0 = 0.0, fh1 = 0.0, fh2 = 0.0;
    if (hllc_core(gp, kEpsWaf, eh, eu, ev, fh, fu, fv, o)) {
        fh0 = o.fh[0]; fh1 = o.fh[1]; fh2 = o.fh[2];
    }
    flux[0] = fh0;
    flux[1] = n1 * fh1 - n2 * fh2;
    flux[2] = n2 * fh1 + n1 * fh2;
}

// Rotation to the face frame with the cell dry threshold
// (flux_riemann.cuf:105-112, :186-193, :240-258).
template <class T>
CF_HD inline void rot(double tolisec, double n1, double n2, T h0, T hu, T hv, T &h, T &un, T &ut) {
    h = h0;
    if (h > tolisec) {
        un = (hu * n1 + hv * n2) / h;
        ut = (hv * n1 - hu * n2) / h;
    } else {
        h = 0.0;
        un = 0.0;
        ut = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Reverse (adjoint) forms.  Each takes the forward inputs (the forward
// intermediates are recomputed), the cotangents of the forward outputs, and
// adds the cotangents of the inputs to the accumulators it is given.

// Reverse of rot for a cell state (h0, hu, hv): cotangents of (h, un, ut).
CF_HD inline void rot_rev(double tolisec, double n1, double n2, double h0, double hu, double hv, double bh,
                          double bun, double but, double &bh0, double &bhu, double &bhv) {
    if (!(h0 > tolisec)) return;  // all outputs are constants
    // derivative-only arithmetic: one reciprocal (no branch depends on it)
    const double inv = 1.0 / h0;
    const double un = (hu * n1 + hv * n2) * inv;
    const double ut = (hv * n1 - hu * n2) * inv;
    bhu += (bun * n1 - but * n2) * inv;
    bhv += (bun * n2 + but * n1) * inv;
    bh0 += bh - (bun * un + but * ut) * inv;
}

// Reverse of hllc_core.  Output cotangents: fh[0..2], f1[0..1], f2[0..1],
// sl, sr, set (the first-order solver uses only fh; WAF uses all of them).
// g[0..5] accumulates the cotangents of (hl, ul, vl, hr, ur, vr).
CF_HD inline void hllc_core_rev(double gp, double eps, double hl, double ul, double vl, double hr, double ur,
                                double vr, const double bfh[3], const double bf1in[2], const double bf2in[2],
                                double bsl, double bsr, double bset, double g[6]) {
    if (hl < eps && hr < eps) return;
    // forward intermediates
    const double al = sqrt(gp * hl), ar = sqrt(gp * hr);
    const double S = hl + hr, A = al + ar;
    const double het = S / 2.0 - (((ur - ul) * S) / A) / 4.0;
    const double uet = (ul + ur) / 2.0 - ((hr - hl) * A) / S;
    const bool pql_var = !(het < hl) && hl > eps;
    const bool pqr_var = !(het < hr) && hr > eps;
    const double pql = (het < hl) ? 1.0 : (pql_var ? sqrt(((0.5 * (het + hl)) * het) / (hl * hl)) : 0.0);
    const double pqr = (het < hr) ? 1.0 : (pqr_var ? sqrt(((0.5 * (het + hr)) * het) / (hr * hr)) : 0.0);
    double sl, sr;
    int sb;  // wave-speed case
    if (hl > eps) {
        if (hr > eps) {
            sl = ul - al * pql;
            sr = ur + ar * pqr;
            sb = 0;
        } else {
            sl = ul - al;
            sr = ul + 2.0 * al;
            sb = 1;
        }
    } else {
        sl = ur - 2.0 * ar;
        sr = ur + ar;
        sb = 2;
    }
    const double denom = hr * (ur - sr) - hl * (ul - sl);
    const bool set_uet = fabs(denom) < eps;
    const double set = set_uet ? uet : ((sl * hr) * (ur - sr) - (sr * hl) * (ul - sl)) / denom;
    const double f1[3] = {hl * ul, hl * (ul * ul) + ((gp * hl) * hl) / 2.0, (hl * ul) * vl};
    const double f2[3] = {hr * ur, hr * (ur * ur) + ((gp * hr) * hr) / 2.0, (hr * ur) * vr};
    const double q1[2] = {hl, hl * ul};
    const double q2[2] = {hr, hr * ur};

    double bhl = 0.0, bul = 0.0, bvl = 0.0, bhr = 0.0, bur = 0.0, bvr = 0.0;
    double bf1[3] = {bf1in[0], bf1in[1], 0.0}, bf2[3] = {bf2in[0], bf2in[1], 0.0};
    double bq1[2] = {0.0, 0.0}, bq2[2] = {0.0, 0.0};
    // fh
    if (sl > 0.0) {
        for (int k = 0; k < 3; k++) bf1[k] += bfh[k];
    } else if (sr < 0.0) {
        for (int k = 0; k < 3; k++) bf2[k] += bfh[k];
    } else {
        const double D = (sr - sl) + eps, invD = 1.0 / D;
        double fh[2];
        for (int k = 0; k < 2; k++) fh[k] = ((sr * f1[k] - sl * f2[k]) + (sl * sr) * (q2[k] - q1[k])) * invD;
        double bfk[2] = {bfh[0], bfh[1]};
        if (set > 0.0) {
            bfk[0] += bfh[2] * vl;
            bvl += bfh[2] * fh[0];
        } else {
            bfk[0] += bfh[2] * vr;
            bvr += bfh[2] * fh[0];
        }
        double bD = 0.0;
        for (int k = 0; k < 2; k++) {
            const double bN = bfk[k] * invD;
            const double dq = q2[k] - q1[k];
            bD -= bN * fh[k];
            bsr += bN * f1[k] + bN * sl * dq;
            bsl += -bN * f2[k] + bN * sr * dq;
            bf1[k] += bN * sr;
            bf2[k] -= bN * sl;
            bq2[k] += bN * (sl * sr);
            bq1[k] -= bN * (sl * sr);
        }
        bsr += bD;
        bsl -= bD;
    }
    // set (used by the WAF weights only)
    double bA = 0.0, bS = 0.0;
    if (bset != 0.0) {
        if (set_uet) {
            const double E = (hr - hl) * A;
            bul += bset / 2.0;
            bur += bset / 2.0;
            const double bE = -bset / S;
            bS += bset * E / (S * S);
            bhr += bE * A;
            bhl -= bE * A;
            bA += bE * (hr - hl);
        } else {
            const double bnum = bset / denom;
            const double bden = -bset * set / denom;
            bsl += bnum * (hr * (ur - sr)) + bnum * (sr * hl);
            bhr += bnum * (sl * (ur - sr));
            bur += bnum * (sl * hr);
            bsr += -bnum * (sl * hr) - bnum * (hl * (ul - sl));
            bhl += -bnum * (sr * (ul - sl));
            bul += -bnum * (sr * hl);
            bhr += bden * (ur - sr);
            bur += bden * hr;
            bsr -= bden * hr;
            bhl -= bden * (ul - sl);
            bul -= bden * hl;
            bsl += bden * hl;
        }
    }
    // wave speeds
    double bal = 0.0, bar = 0.0, bpql = 0.0, bpqr = 0.0;
    if (sb == 0) {
        bul += bsl;
        bal -= bsl * pql;
        bpql -= bsl * al;
        bur += bsr;
        bar += bsr * pqr;
        bpqr += bsr * ar;
    } else if (sb == 1) {
        bul += bsl + bsr;
        bal += -bsl + 2.0 * bsr;
    } else {
        bur += bsl + bsr;
        bar += -2.0 * bsl + bsr;
    }
    // pql, pqr: sqrt(Q), Q = (0.5 (het + h)) het / h^2
    double bhet = 0.0;
    if (pql_var && bpql != 0.0) {
        const double H2 = hl * hl, Q = ((0.5 * (het + hl)) * het) / H2;
        const double bP = (bpql / (2.0 * pql)) / H2;
        bhet += bP * (het + 0.5 * hl);
        bhl += bP * (0.5 * het) - (bP * Q) * (2.0 * hl);
    }
    if (pqr_var && bpqr != 0.0) {
        const double H2 = hr * hr, Q = ((0.5 * (het + hr)) * het) / H2;
        const double bP = (bpqr / (2.0 * pqr)) / H2;
        bhet += bP * (het + 0.5 * hr);
        bhr += bP * (0.5 * het) - (bP * Q) * (2.0 * hr);
    }
    // het = S/2 - ((ur - ul) S / A) / 4
    if (bhet != 0.0) {
        const double invA = 1.0 / A;
        const double W = ((ur - ul) * S) * invA;
        const double bW = -bhet * 0.25;
        bS += bhet * 0.5 + bW * ((ur - ul) * invA);
        bur += bW * (S * invA);
        bul -= bW * (S * invA);
        bA -= bW * (W * invA);
    }
    bhl += bS;
    bhr += bS;
    bal += bA;
    bar += bA;
    if (al > 0.0) bhl += bal * (gp / (2.0 * al));
    if (ar > 0.0) bhr += bar * (gp / (2.0 * ar));
    // f1, f2, q1, q2
    bhl += bf1[0] * ul + bf1[1] * (ul * ul + gp * hl) + bf1[2] * (ul * vl) + bq1[0] + bq1[1] * ul;
    bul += bf1[0] * hl + bf1[1] * (2.0 * hl * ul) + bf1[2] * (hl * vl) + bq1[1] * hl;
    bvl += bf1[2] * (hl * ul);
    bhr += bf2[0] * ur + bf2[1] * (ur * ur + gp * hr) + bf2[2] * (ur * vr) + bq2[0] + bq2[1] * ur;
    bur += bf2[0] * hr + bf2[1] * (2.0 * hr * ur) + bf2[2] * (hr * vr) + bq2[1] * hr;
    bvr += bf2[2] * (hr * ur);
    g[0] += bhl;
    g[1] += bul;
    g[2] += bvl;
    g[3] += bhr;
    g[4] += bur;
    g[5] += bvr;
}

// Cotangent of a global-frame flux (f0, n1 f1 - n2 f2, n2 f1 + n1 f2) mapped
// to the face-frame components.
CF_HD inline void unrotate_rev(double n1, double n2, const double bflux[3], double b[3]) {
    b[0] = bflux[0];
    b[1] = n1 * bflux[1] + n2 * bflux[2];
    b[2] = -n2 * bflux[1] + n1 * bflux[2];
}

// Reverse of riadh_flux: g[0..5] += cotangents of (hl, ul, vl, hr, ur, vr).
// The region's flux is reversed here; the wave speeds S_L, S_R, S* (the
// same estimates as hllc_core with eps = 1e-6f) through hllc_core_rev.
CF_HD inline void riadh_flux_rev(double gp, double n1, double n2, double hl, double ul, double vl, double hr,
                                 double ur, double vr, const double bflux[3], double g[6]) {
    const double eps = kEpsRiadh;
    if (hl < eps && hr < eps) return;
    const double al = sqrt(gp * hl), ar = sqrt(gp * hr);
    const double het = (hl + hr) / 2.0 - (((ur - ul) * (hl + hr)) / (al + ar)) / 4.0;
    const double uet = (ul + ur) / 2.0 - ((hr - hl) * (al + ar)) / (hl + hr);
    double pql, pqr, sl, sr, set;
    if (het < hl) pql = 1.0;
    else pql = (hl > eps) ? sqrt(((0.5 * (het + hl)) * het) / (hl * hl)) : 0.0;
    if (het < hr) pqr = 1.0;
    else pqr = (hr > eps) ? sqrt(((0.5 * (het + hr)) * het) / (hr * hr)) : 0.0;
    if (hl > eps) {
        sl = ul - al * pql;
        if (hr > eps) {
            sr = ur + ar * pqr;
        } else {
            sl = ul - al;
            sr = ul + 2.0 * al;
        }
    } else {
        sl = ur - 2.0 * ar;
        sr = ur + ar;
    }
    const double denom = hr * (ur - sr) - hl * (ul - sl);
    set = (fabs(denom) < eps) ? uet : ((sl * hr) * (ur - sr) - (sr * hl) * (ul - sl)) / denom;
    double b[3];
    unrotate_rev(n1, n2, bflux, b);
    double bh = 0.0, bu = 0.0, bv = 0.0, bsk = 0.0, bset = 0.0;  // side K of the selected region
    int side;  // 0 left, 1 right
    bool star;
    if (0.0 < sl) {
        side = 0;
        star = false;
    } else if (0.0 < set && 0.0 > sl) {
        side = 0;
        star = true;
    } else if (0.0 > set && 0.0 < sr) {
        side = 1;
        star = true;
    } else {
        side = 1;
        star = false;
    }
    const double h = side ? hr : hl, u = side ? ur : ul, v = side ? vr : vl, sk = side ? sr : sl;
    // physical flux F(U_K) = (h u, h u^2 + g h^2 / 2, h u v)
    bh += b[0] * u + b[1] * (u * u + gp * h) + b[2] * (u * v);
    bu += b[0] * h + b[1] * (2.0 * h * u) + b[2] * (h * v);
    bv += b[2] * (h * u);
    if (star) {
        // + S_K (P (1, S*, v) - (h, h u, h v)), P = h (S_K - u) / (S_K - S*)
        const bool pa = fabs(sk - set) > eps;
        const double P = pa ? (h * (sk - u)) / (sk - set) : 0.0;
        bsk += b[0] * (P - h) + b[1] * (P * set - h * u) + b[2] * (P * v - h * v);
        const double bP = b[0] * sk + b[1] * (sk * set) + b[2] * (sk * v);
        bset += b[1] * (sk * P);
        bh -= b[0] * sk + b[1] * (sk * u) + b[2] * (sk * v);
        bu -= b[1] * (sk * h);
        bv += b[2] * (sk * (P - h));
        if (pa) {
            const double inv = 1.0 / (sk - set);
            bh += bP * ((sk - u) * inv);
            bu -= bP * (h * inv);
            bsk += bP * (h * inv) - bP * (P * inv);
            bset += bP * (P * inv);
        }
    }
    // constant indices (a runtime index would keep g in private memory)
    if (side) {
        g[3] += bh;
        g[4] += bu;
        g[5] += bv;
    } else {
        g[0] += bh;
        g[1] += bu;
        g[2] += bv;
    }
    const double z3[3] = {0.0, 0.0, 0.0}, z2[2] = {0.0, 0.0};
    hllc_core_rev(gp, eps, hl, ul, vl, hr, ur, vr, z3, z2, z2, side ? 0.0 : bsk, side ? bsk : 0.0, bset, g);
}

// Reverse of hllc_flux: g[0..5] += cotangents of (hl, ul, vl, hr, ur, vr).
CF_HD inline void hllc_flux_rev(double gp, double n1, double n2, double hl, double ul, double vl, double hr,
                                double ur, double vr, const double bflux[3], double g[6]) {
    double bfh[3];
    unrotate_rev(n1, n2, bflux, bfh);
    const double z2[2] = {0.0, 0.0};
    hllc_core_rev(gp, kEpsHllc, hl, ul, vl, hr, ur, vr, bfh, z2, z2, 0.0, 0.0, 0.0, g);
}

// Value and partial derivatives (with respect to r and c) of limiter().
CF_HD inline double limiter_d(int ilim, double r, double c, double &dr, double &dc) {
    const double ac = fabs(c);
    // d|c|/dc with the dual type's convention at c = 0 (+1)
    const double sc = c < 0.0 ? -1.0 : 1.0;
    double dac = 0.0;
    double v;
    dr = 0.0;
    switch (ilim) {
    case 1:
        if (r <= 0.0) v = 1.0;
        else if (r > 0.0 && r < 0.5) { v = 1.0 - (2.0 * (1.0 - ac)) * r; dr = -(2.0 * (1.0 - ac)); dac = 2.0 * r; }
        else if (r > 0.5 && r < 1.0) { v = ac; dac = 1.0; }
        else if (r > 1.0 && r < 2.0) { v = 1.0 - (1.0 - ac) * r; dr = -(1.0 - ac); dac = r; }
        else { v = 2.0 * ac - 1.0; dac = 2.0; }
        break;
    case 2:
        if (r <= 0.0) v = 1.0;
        else {
            v = 1.0 - ((2.0 * (1.0 - ac)) * r) / (1.0 + r);
            dr = -(2.0 * (1.0 - ac)) / ((1.0 + r) * (1.0 + r));
            dac = (2.0 * r) / (1.0 + r);
        }
        break;
    case 3:
        if (r <= 0.0) v = 1.0;
        else {
            const double q = (r * (1.0 + r)) / (1.0 + r * r);
            v = 1.0 - (((1.0 - ac) * r) * (1.0 + r)) / (1.0 + r * r);
            dr = -(1.0 - ac) * ((1.0 + 2.0 * r - r * r) / ((1.0 + r * r) * (1.0 + r * r)));
            dac = q;
        }
        break;
    case 4:
        if (r <= 0.0) v = 1.0;
        else if (r <= 1.0) { v = 1.0 - (1.0 - ac) * r; dr = -(1.0 - ac); dac = r; }
        else { v = ac; dac = 1.0; }
        break;
    case 0:
    default: v = 1.0; break;
    }
    dc = dac * sc;
    return v;
}

// Reverse of waf_flux.  g[0..9] += cotangents of
// (hl, ul, vl, hr, ur, vr, hl_upw, vl_upw, hr_upw, vr_upw).
CF_HD inline void waf_flux_rev(double gp, int ilim, double n1, double n2, double hl, double ul, double vl, double hr,
                               double ur, double vr, double hl_upw, double vl_upw, double hr_upw, double vr_upw,
                               double d_mrl, double d_mkvol, double d_mkvoisin, double dt, const double bflux[3],
                               double g[10], bool freeze_lim = false) {
    const double eps = kEpsWaf;
    HllcT<double> o;
    if (!hllc_core(gp, eps, hl, ul, vl, hr, ur, vr, o)) return;
    double bfw[3];
    unrotate_rev(n1, n2, bflux, bfw);
    const double dtoverdx = dt / d_mrl;
    const double cr = o.sr * dtoverdx, cl = o.sl * dtoverdx, cet = o.set * dtoverdx;
    const double dlr = (hr - hl) / d_mrl, dll = (hr - hl) / d_mrl, dlet = (vr - vl) / d_mrl;
    const double dur = (o.sr > 0.0) ? (hl - hl_upw) / d_mkvol : (hr_upw - hr) / d_mkvoisin;
    const double dul = (o.sl > 0.0) ? (hl - hl_upw) / d_mkvol : (hr_upw - hr) / d_mkvoisin;
    const double duet = (o.set > 0.0) ? (vl - vl_upw) / d_mkvol : (vr_upw - vr) / d_mkvoisin;
    const double rl = dul / (dll + eps), rr = dur / (dlr + eps), ret = duet / (dlet + eps);
    double drl, dcl, drr, dcr, dret, dcet;
    const double lim_l = limiter_d(ilim, rl, cl, drl, dcl);
    const double lim_r = limiter_d(ilim, rr, cr, drr, dcr);
    const double lim_et = limiter_d(ilim, ret, cet, dret, dcet);
    const double sgl = copysign(1.0, o.sl), sgr = copysign(1.0, o.sr), sge = copysign(1.0, o.set);
    const double wl = (1.0 + sgl * lim_l) / 2.0;
    const double wlet = (1.0 + sge * lim_et) / 2.0;
    const double wlr = (sgr * lim_r - sgl * lim_l) / 2.0;
    const double wr = (1.0 - sgr * lim_r) / 2.0;
    const double wret = (1.0 - sge * lim_et) / 2.0;
    const double fw0 = (wl * o.f1[0] + wlr * o.fh[0]) + wr * o.f2[0];

    double bhl = 0.0, bul = 0.0, bvl = 0.0, bhr = 0.0, bur = 0.0, bvr = 0.0;
    double bhlu = 0.0, bvlu = 0.0, bhru = 0.0, bvru = 0.0;
    // fw2 = (wlet vl + wret vr) fw0
    const double V = wlet * vl + wret * vr;
    const double bV = bfw[2] * fw0;
    const double bfw0 = bfw[0] + bfw[2] * V;
    const double bwlet = bV * vl, bwret = bV * vr;
    bvl += bV * wlet;
    bvr += bV * wret;
    // fw0, fw1
    const double bwl = bfw0 * o.f1[0] + bfw[1] * o.f1[1];
    const double bwlr = bfw0 * o.fh[0] + bfw[1] * o.fh[1];
    const double bwr = bfw0 * o.f2[0] + bfw[1] * o.f2[1];
    const double bf1[2] = {bfw0 * wl, bfw[1] * wl};
    const double bfh[3] = {bfw0 * wlr, bfw[1] * wlr, 0.0};
    const double bf2[2] = {bfw0 * wr, bfw[1] * wr};
    // weights -> limiter values
    const double blim_l = (bwl * sgl - bwlr * sgl) / 2.0;
    const double blim_r = (bwlr * sgr - bwr * sgr) / 2.0;
    const double blim_et = (bwlet * sge - bwret * sge) / 2.0;
    // limiters -> (r, c); nothing when the limiter values are frozen
    const double fz = freeze_lim ? 0.0 : 1.0;
    const double brl = fz * blim_l * drl, brr = fz * blim_r * drr, bret = fz * blim_et * dret;
    const double bcl = fz * blim_l * dcl, bcr = fz * blim_r * dcr, bcet = fz * blim_et * dcet;
    // r = du / (dl + eps)
    const double bdul = brl / (dll + eps), bdur = brr / (dlr + eps), bduet = bret / (dlet + eps);
    const double bdll = -brl * (rl / (dll + eps)), bdlr = -brr * (rr / (dlr + eps));
    const double bdlet = -bret * (ret / (dlet + eps));
    bhr += (bdll + bdlr) / d_mrl;
    bhl -= (bdll + bdlr) / d_mrl;
    bvr += bdlet / d_mrl;
    bvl -= bdlet / d_mrl;
    if (o.sr > 0.0) {
        bhl += bdur / d_mkvol;
        bhlu -= bdur / d_mkvol;
    } else {
        bhru += bdur / d_mkvoisin;
        bhr -= bdur / d_mkvoisin;
    }
    if (o.sl > 0.0) {
        bhl += bdul / d_mkvol;
        bhlu -= bdul / d_mkvol;
    } else {
        bhru += bdul / d_mkvoisin;
        bhr -= bdul / d_mkvoisin;
    }
    if (o.set > 0.0) {
        bvl += bduet / d_mkvol;
        bvlu -= bduet / d_mkvol;
    } else {
        bvru += bduet / d_mkvoisin;
        bvr -= bduet / d_mkvoisin;
    }
    double gg[6] = {bhl, bul, bvl, bhr, bur, bvr};
    hllc_core_rev(gp, eps, hl, ul, vl, hr, ur, vr, bfh, bf1, bf2, bcl * dtoverdx, bcr * dtoverdx, bcet * dtoverdx,
                  gg);
    for (int k = 0; k < 6; k++) g[k] += gg[k];
    g[6] += bhlu;
    g[7] += bvlu;
    g[8] += bhru;
    g[9] += bvru;
}

// slope_limiter with its partial derivatives, the branch and the argument
// of each max/min taken as the forward takes them (dmax(a, b) keeps b on a
// tie, as the dual type does); d|p|/dp = +1 at p = 0 (the dual type's vabs).
CF_HD inline double slope_limiter_d(int ilim, double r1, double r2, double &d1, double &d2) {
    d1 = 0.0;
    d2 = 0.0;
    switch (ilim) {
    case 1: {
        double a, da1, da2, b, db1, db2;
        if (r2 > 0.0) {
            if (2.0 * r1 < r2) { a = 2.0 * r1; da1 = 2.0; da2 = 0.0; } else { a = r2; da1 = 0.0; da2 = 1.0; }
            if (r1 < 2.0 * r2) { b = r1; db1 = 1.0; db2 = 0.0; } else { b = 2.0 * r2; db1 = 0.0; db2 = 2.0; }
            double m, dm1, dm2;
            if (0.0 > a) { m = 0.0; dm1 = 0.0; dm2 = 0.0; } else { m = a; dm1 = da1; dm2 = da2; }
            if (m > b) { d1 = dm1; d2 = dm2; return m; }
            d1 = db1;
            d2 = db2;
            return b;
        }
        if (2.0 * r1 > r2) { a = 2.0 * r1; da1 = 2.0; da2 = 0.0; } else { a = r2; da1 = 0.0; da2 = 1.0; }
        if (r1 > 2.0 * r2) { b = r1; db1 = 1.0; db2 = 0.0; } else { b = 2.0 * r2; db1 = 0.0; db2 = 2.0; }
        double m, dm1, dm2;
        if (0.0 < a) { m = 0.0; dm1 = 0.0; dm2 = 0.0; } else { m = a; dm1 = da1; dm2 = da2; }
        if (m < b) { d1 = dm1; d2 = dm2; return m; }
        d1 = db1;
        d2 = db2;
        return b;
    }
    case 2: {
        const double p = r1 * r2, sp = p < 0.0 ? -1.0 : 1.0;
        const double num = p + fabs(p), den = (r1 + r2) + kEpsLeer;
        const double v = num / den;
        d1 = (r2 * (1.0 + sp) - v) / den;
        d2 = (r1 * (1.0 + sp) - v) / den;
        return v;
    }
    case 3:
        if (r2 > 0.0) {
            const double in = r1 < r2 ? r1 : r2;
            if (0.0 > in) return 0.0;
            if (r1 < r2) d1 = 1.0;
            else d2 = 1.0;
            return in;
        } else {
            const double in = r1 > r2 ? r1 : r2;
            if (0.0 < in) return 0.0;
            if (r1 > r2) d1 = 1.0;
            else d2 = 1.0;
            return in;
        }
    default: return 0.0;
    }
}

// Reverse of muscl_flux (dt, the limiter branches, the dry-velocity
// switches and the clamps at zero as taken in the forward).  g[0..11] +=
// cotangents of (hl, ul, vl, hr, ur, vr, hl_upw, ul_upw, vl_upw, hr_upw,
// ur_upw, vr_upw).
CF_HD inline void muscl_flux_rev(double gp, int ilim, double tolisec, double n1, double n2, double hl, double ul,
                                 double vl, double hr, double ur, double vr, double hlu, double ulu, double vlu,
                                 double hru, double uru, double vru, double d_mrl, double d_mkvol,
                                 double d_mkvoisin, double dt, const double bflux[3], double g[12]) {
    const double d = d_mrl, dk = d_mkvol, dv = d_mkvoisin;
    const double dtoverdx = dt / d_mrl, hdt = 0.5 * dtoverdx, hg = 0.5 * gp;
    // forward: reconstruction I
    double sLh1, sLh2, sLu1 = 0.0, sLu2 = 0.0, sLv1 = 0.0, sLv2 = 0.0;
    const double dhL = slope_limiter_d(ilim, (hl - hlu) / dk, (hr - hl) / d, sLh1, sLh2);
    const double rLl = hl - (0.5 * dhL) * d, rLr = hl + (0.5 * dhL) * d;
    const double hll = dmax(0.0, rLl), hlr = dmax(0.0, rLr);
    const bool velL = !(hlr < tolisec || hll < tolisec);
    double ull = 0.0, ulr = 0.0, vll = 0.0, vlr = 0.0;
    if (velL) {
        const double du = slope_limiter_d(ilim, (ul - ulu) / dk, (ur - ul) / d, sLu1, sLu2);
        const double dw = slope_limiter_d(ilim, (vl - vlu) / dk, (vr - vl) / d, sLv1, sLv2);
        ull = ul - (0.5 * du) * d;
        vll = vl - (0.5 * dw) * d;
        ulr = ul + (0.5 * du) * d;
        vlr = vl + (0.5 * dw) * d;
    }
    double sRh1, sRh2, sRu1 = 0.0, sRu2 = 0.0, sRv1 = 0.0, sRv2 = 0.0;
    const double dhR = slope_limiter_d(ilim, (hr - hl) / d, (hru - hr) / dv, sRh1, sRh2);
    const double rRr = hr + (0.5 * dhR) * d, rRl = hr - (0.5 * dhR) * d;
    const double hrr = dmax(0.0, rRr), hrl = dmax(0.0, rRl);
    const bool velR = !(hrl < tolisec || hrr < tolisec);
    double url = 0.0, urr = 0.0, vrl = 0.0, vrr = 0.0;
    if (velR) {
        const double du = slope_limiter_d(ilim, (ur - ul) / d, (uru - ur) / dv, sRu1, sRu2);
        const double dw = slope_limiter_d(ilim, (vr - vl) / d, (vru - vr) / dv, sRv1, sRv2);
        urr = ur + (0.5 * du) * d;
        vrr = vr + (0.5 * dw) * d;
        url = ur - (0.5 * du) * d;
        vrl = vr - (0.5 * dw) * d;
    }
    // forward: reconstruction II
    const double eh = hlr + hdt * (hll * ull - hlr * ulr);
    const double eu = ulr + hdt * ((hll * (ull * ull) + hg * (hll * hll)) - (hlr * (ulr * ulr) + hg * (hlr * hlr)));
    const double ev = vlr + hdt * ((hll * ull) * vll - (hlr * ulr) * vlr);
    const double fh = hrl + hdt * (hrl * url - hrr * urr);
    const double fu = url + hdt * ((hrl * (url * url) + hg * (hrl * hrl)) - (hrr * (urr * urr) + hg * (hrr * hrr)));
    const double fv = vrl + hdt * ((hrl * url) * vrl - (hrr * urr) * vrr);
    if (eh < kEpsWaf && fh < kEpsWaf) return;  // dry/dry: zero flux
    // reverse: the HLLC solver on the evolved values
    double bfh[3];
    unrotate_rev(n1, n2, bflux, bfh);
    const double z2[2] = {0.0, 0.0};
    double ge[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    hllc_core_rev(gp, kEpsWaf, eh, eu, ev, fh, fu, fv, bfh, z2, z2, 0.0, 0.0, 0.0, ge);
    const double beh = ge[0], beu = ge[1], bev = ge[2], bfhh = ge[3], bfu = ge[4], bfv = ge[5];
    // reverse: reconstruction II
    double bhll = 0.0, bhlr = 0.0, bull = 0.0, bulr = 0.0, bvll = 0.0, bvlr = 0.0;
    bhlr += beh * (1.0 - hdt * ulr);
    bhll += beh * hdt * ull;
    bull += beh * hdt * hll;
    bulr -= beh * hdt * hlr;
    bulr += beu * (1.0 - hdt * (2.0 * hlr * ulr));
    bhll += beu * hdt * (ull * ull + 2.0 * hg * hll);
    bull += beu * hdt * (2.0 * hll * ull);
    bhlr -= beu * hdt * (ulr * ulr + 2.0 * hg * hlr);
    bvlr += bev * (1.0 - hdt * (hlr * ulr));
    bhll += bev * hdt * (ull * vll);
    bull += bev * hdt * (hll * vll);
    bvll += bev * hdt * (hll * ull);
    bhlr -= bev * hdt * (ulr * vlr);
    bulr -= bev * hdt * (hlr * vlr);
    double bhrl = 0.0, bhrr = 0.0, burl = 0.0, burr = 0.0, bvrl = 0.0, bvrr = 0.0;
    bhrl += bfhh * (1.0 + hdt * url);
    burl += bfhh * hdt * hrl;
    bhrr -= bfhh * hdt * urr;
    burr -= bfhh * hdt * hrr;
    burl += bfu * (1.0 + hdt * (2.0 * hrl * url));
    bhrl += bfu * hdt * (url * url + 2.0 * hg * hrl);
    bhrr -= bfu * hdt * (urr * urr + 2.0 * hg * hrr);
    burr -= bfu * hdt * (2.0 * hrr * urr);
    bvrl += bfv * (1.0 + hdt * (hrl * url));
    bhrl += bfv * hdt * (url * vrl);
    burl += bfv * hdt * (hrl * vrl);
    bhrr -= bfv * hdt * (urr * vrr);
    burr -= bfv * hdt * (hrr * vrr);
    bvrr -= bfv * hdt * (hrr * urr);
    // reverse: reconstruction I
    double bhl = 0.0, bul = 0.0, bvl = 0.0, bhr = 0.0, bur = 0.0, bvr = 0.0;
    double bhlu = 0.0, bulu = 0.0, bvlu = 0.0, bhru = 0.0, buru = 0.0, bvru = 0.0;
    if (velL) {
        bul += bull + bulr;
        bvl += bvll + bvlr;
        const double bdu = 0.5 * d * (bulr - bull), bdw = 0.5 * d * (bvlr - bvll);
        const double a1 = bdu * sLu1, a2 = bdu * sLu2, c1 = bdw * sLv1, c2 = bdw * sLv2;
        bul += a1 / dk - a2 / d;
        bulu -= a1 / dk;
        bur += a2 / d;
        bvl += c1 / dk - c2 / d;
        bvlu -= c1 / dk;
        bvr += c2 / d;
    }
    {
        const double bl = !(0.0 > rLl) ? bhll : 0.0, br = !(0.0 > rLr) ? bhlr : 0.0;
        bhl += bl + br;
        const double bdh = 0.5 * d * (br - bl);
        const double a1 = bdh * sLh1, a2 = bdh * sLh2;
        bhl += a1 / dk - a2 / d;
        bhlu -= a1 / dk;
        bhr += a2 / d;
    }
    if (velR) {
        bur += burr + burl;
        bvr += bvrr + bvrl;
        const double bdu = 0.5 * d * (burr - burl), bdw = 0.5 * d * (bvrr - bvrl);
        const double a1 = bdu * sRu1, a2 = bdu * sRu2, c1 = bdw * sRv1, c2 = bdw * sRv2;
        bur += a1 / d - a2 / dv;
        bul -= a1 / d;
        buru += a2 / dv;
        bvr += c1 / d - c2 / dv;
        bvl -= c1 / d;
        bvru += c2 / dv;
    }
    {
        const double brr = !(0.0 > rRr) ? bhrr : 0.0, brl = !(0.0 > rRl) ? bhrl : 0.0;
        bhr += brr + brl;
        const double bdh = 0.5 * d * (brr - brl);
        const double a1 = bdh * sRh1, a2 = bdh * sRh2;
        bhr += a1 / d - a2 / dv;
        bhl -= a1 / d;
        bhru += a2 / dv;
    }
    g[0] += bhl;
    g[1] += bul;
    g[2] += bvl;
    g[3] += bhr;
    g[4] += bur;
    g[5] += bvr;
    g[6] += bhlu;
    g[7] += bulu;
    g[8] += bvlu;
    g[9] += bhru;
    g[10] += buru;
    g[11] += bvru;
}

}  // namespace cf
