// Device-side numerics of the St-Venant step: project pow, HLLC and WAF
// Riemann solvers, limiter.  Every expression keeps the operation order of
// the reference Fortran (flux_riemann.cuf, source_terms.cuf) so that the
// results are bit-identical to a CPU evaluation of the same formulas; the
// translation unit is compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace cf {

// Default-real literals of the reference are single precision, widened.
constexpr double kEpsWaf = (double)1e-16f;   // flux_riemann.cuf:749
constexpr double kEpsHllc = (double)1e-8f;   // flux_riemann.cuf:926
constexpr double kDtLow = (double)1e-12f;    // cfl_condition.cuf:129
constexpr double kCote = (double)1.8f;       // maillage.cuf:157

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

__device__ __forceinline__ void two_prod(double a, double b, double &p, double &e) {
    const double split = 134217729.0;  // 2^27 + 1 (Veltkamp)
    p = a * b;
    double t = split * a;
    double ah = t - (t - a), al = a - ah;
    t = split * b;
    double bh = t - (t - b), bl = b - bh;
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// x^y for x > 0: double-double log (atanh series) and Cody-Waite exp.
// Uses only +,-,*,/, floor and bit operations: same bits as the host build.
__device__ inline double cf_pow(double x, double y) {
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

// Wave speeds and HLLC flux in the face frame (flux_riemann.cuf:930-1020 for
// HLLC, :753-845 for the first half of WAF).  Returns false for a dry/dry
// face: the reference leaves al, ar, pql, pqr unset there; the flux is zero.
struct Hllc {
    double f1[3], f2[3], fh[3];
    double sl, sr, set;
};

__device__ __forceinline__ bool hllc_core(double gp, double eps, double hl, double ul,
                                          double vl, double hr, double ur, double vr,
                                          Hllc &o) {
    if (hl < eps && hr < eps) return false;
    double al = sqrt(gp * hl);
    double ar = sqrt(gp * hr);
    double het = (hl + hr) / 2.0 - (((ur - ul) * (hl + hr)) / (al + ar)) / 4.0;
    double uet = (ul + ur) / 2.0 - ((hr - hl) * (al + ar)) / (hl + hr);
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
    double denom = hr * (ur - sr) - hl * (ul - sl);
    if (fabs(denom) < eps) set = uet;
    else set = ((sl * hr) * (ur - sr) - (sr * hl) * (ul - sl)) / denom;
    double q1[3] = {hl, hl * ul, hl * vl};
    double q2[3] = {hr, hr * ur, hr * vr};
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
__device__ __forceinline__ void hllc_flux(double gp, double n1, double n2, double hl,
                                          double ul, double vl, double hr, double ur,
                                          double vr, double flux[3]) {
    Hllc o;
    double fh0 = 0.0, fh1 = 0.0, fh2 = 0.0;
    if (hllc_core(gp, kEpsHllc, hl, ul, vl, hr, ur, vr, o)) {
        fh0 = o.fh[0]; fh1 = o.fh[1]; fh2 = o.fh[2];
    }
    flux[0] = fh0;
    flux[1] = n1 * fh1 - n2 * fh2;
    flux[2] = n2 * fh1 + n1 * fh2;
}

// limiter (flux_riemann.cuf:465-524); unknown ilimiteur -> 1.
__device__ __forceinline__ double limiter(int ilim, double r, double c) {
    double ac = fabs(c);
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
__device__ __forceinline__ void waf_flux(double gp, int ilim, double n1, double n2,
                                         double hl, double ul, double vl, double hr,
                                         double ur, double vr, double hl_upw,
                                         double vl_upw, double hr_upw, double vr_upw,
                                         double d_mrl, double d_mkvol, double d_mkvoisin,
                                         double dt, double flux[3]) {
    const double eps = kEpsWaf;
    Hllc o;
    double fw0 = 0.0, fw1 = 0.0, fw2 = 0.0;
    if (hllc_core(gp, eps, hl, ul, vl, hr, ur, vr, o)) {
        double dtoverdx = dt / d_mrl;
        double cr = o.sr * dtoverdx;
        double cl = o.sl * dtoverdx;
        double cet = o.set * dtoverdx;
        double dlr = (hr - hl) / d_mrl;
        double dll = (hr - hl) / d_mrl;
        double dlet = (vr - vl) / d_mrl;
        double dur = (o.sr > 0.0) ? (hl - hl_upw) / d_mkvol : (hr_upw - hr) / d_mkvoisin;
        double dul = (o.sl > 0.0) ? (hl - hl_upw) / d_mkvol : (hr_upw - hr) / d_mkvoisin;
        double duet = (o.set > 0.0) ? (vl - vl_upw) / d_mkvol : (vr_upw - vr) / d_mkvoisin;
        double rl = dul / (dll + eps);
        double rr = dur / (dlr + eps);
        double ret = duet / (dlet + eps);
        double lim_l = limiter(ilim, rl, cl);
        double lim_r = limiter(ilim, rr, cr);
        double lim_et = limiter(ilim, ret, cet);
        double sgl = copysign(1.0, o.sl), sgr = copysign(1.0, o.sr), sge = copysign(1.0, o.set);
        double wl = (1.0 + sgl * lim_l) / 2.0;
        double wlet = (1.0 + sge * lim_et) / 2.0;
        double wlr = (sgr * lim_r - sgl * lim_l) / 2.0;
        double wr = (1.0 - sgr * lim_r) / 2.0;
        double wret = (1.0 - sge * lim_et) / 2.0;
        fw0 = (wl * o.f1[0] + wlr * o.fh[0]) + wr * o.f2[0];
        fw1 = (wl * o.f1[1] + wlr * o.fh[1]) + wr * o.f2[1];
        fw2 = (wlet * vl + wret * vr) * fw0;
    }
    flux[0] = fw0;
    flux[1] = n1 * fw1 - n2 * fw2;
    flux[2] = n2 * fw1 + n1 * fw2;
}

// Rotation to the face frame with the cell dry threshold
// (flux_riemann.cuf:105-112, :186-193, :240-258).
__device__ __forceinline__ void rot(double tolisec, double n1, double n2, double h0,
                                    double hu, double hv, double &h, double &un,
                                    double &ut) {
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

}  // namespace cf
