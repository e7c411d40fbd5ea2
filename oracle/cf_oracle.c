/*
 * CPU oracle for the St-Venant finite-volume time step (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the reference CUDA Fortran hot path of CuteFlow.
 * It is the parity checker: only tests/, the smoke check and bench.py's CPU
 * baseline leg load it; the product never does.  Each function cites the
 * reference file:line it restates (paths relative to the reference root).
 *
 * Arithmetic rules shared with the HIP kernels (and with the reference's
 * Fortran semantics):
 *   - no FMA contraction (built with -ffp-contract=off), no fast-math;
 *   - operation order follows Fortran's left-to-right evaluation, e.g.
 *     gp*h**2/2. is (gp*(h*h))/2 and x**2 is x*x;
 *   - default-real (single precision) literals are widened to double, e.g.
 *     eps=1e-8 in the HLLC solver is (double)1e-8f;
 *   - x**y with a real exponent uses or_pow(), a project-written pow whose
 *     algorithm is repeated bit-for-bit in the device code.
 *
 * Reference undefined behaviour is given a defined meaning:
 *   - dry/dry faces (flux_riemann.cuf:753,930,1055 `goto 20` skips
 *     al,ar,pql,pqr): the Riemann solvers return a zero flux;
 *   - WAF second-neighbour distance d_kvol_v1/v2 used uninitialised
 *     (flux_riemann.cuf:320-331): computed from centroids like :347,355;
 *   - tjv unset when no face of the neighbour matches (flux_riemann.cuf:262-277):
 *     tjv = 1;
 *   - limiter with an unknown ilimiteur leaves A unset: A = 1;
 *   - non-manifold edges (maillage.cuf:904-911 last writer wins): the highest
 *     matching cell id wins;
 *   - duplicated ids in a boundary node list (maillage.cuf:1137-1143 races):
 *     the inlet/outlet number of the first listed node of the edge is used.
 */
#include "cf_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* single-precision literals of the reference, widened to double */
#define LIT_1EM16 ((double)1e-16f) /* flux_riemann.cuf:749 */
#define LIT_1EM8 ((double)1e-8f)   /* flux_riemann.cuf:926 */
#define LIT_1EM6 ((double)1e-6f)   /* flux_riemann.cuf:1051 */
#define LIT_1EM12 ((double)1e-12f) /* cfl_condition.cuf:129, flux_riemann.cuf:453 */
#define LIT_1P8 ((double)1.8f)     /* maillage.cuf:157 */

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* Project pow: x^y = exp(y*log(x)) for x > 0, evaluated with a double-double
 * logarithm (atanh series, Abramowitz & Stegun 4.1.27) and a Cody-Waite
 * reduced exponential (Taylor series), Dekker/Veltkamp exact products.
 * Only +,-,*,/, floor and bit manipulation are used, so the same code gives
 * the same bits on the host and on gfx950.                                  */

typedef union { double d; uint64_t u; } or_bits;

static inline void two_prod(double a, double b, double *p, double *e)
{
    const double split = 134217729.0; /* 2^27 + 1 */
    double t, ah, al, bh, bl;
    *p = a * b;
    t = split * a; ah = t - (t - a); al = a - ah;
    t = split * b; bh = t - (t - b); bl = b - bh;
    *e = ((ah * bh - *p) + ah * bl + al * bh) + al * bl;
}

static inline void two_sum(double a, double b, double *s, double *e)
{
    double bb;
    *s = a + b;
    bb = *s - a;
    *e = (a - (*s - bb)) + (b - bb);
}

double or_pow(double x, double y)
{
    const double ln2_hi = 0x1.62e42ffp-1;
    const double ln2_lo = -0x1.718432a1b0e26p-35;
    const double inv_ln2 = 0x1.71547652b82fep+0;
    const double sqrt2 = 0x1.6a09e667f3bcdp+0;
    or_bits b;
    int e, k;
    double m, f, d, dlo, s, p, plo, r, slo, s2, t, lh, ll, th, tl, tmp, q;
    double rh, ex, scale;

    if (!(x > 0.0) || x != x || y != y) return NAN;
    if (x == INFINITY) return y > 0 ? INFINITY : (y < 0 ? 0.0 : 1.0);

    /* x = 2^e * m with m in [sqrt(1/2), sqrt(2)) */
    b.d = x;
    e = 0;
    if ((b.u >> 52) == 0) { b.d = x * 0x1p54; e = -54; }
    e += (int)((b.u >> 52) & 0x7ff) - 1023;
    b.u = (b.u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    m = b.d;
    if (m >= sqrt2) { m = m * 0.5; e += 1; }

    /* log(m) = 2*atanh(s), s = f/(2+f), f = m-1 (exact) */
    f = m - 1.0;
    d = 2.0 + f;
    dlo = (2.0 - d) + f;
    s = f / d;
    two_prod(s, d, &p, &plo);
    r = ((f - p) - plo) - s * dlo;
    slo = r / d;
    s2 = s * s;
    t = 0x1.47ae147ae147bp-5;            /* 1/25 */
    t = t * s2 + 0x1.642c8590b2164p-5;   /* 1/23 */
    t = t * s2 + 0x1.8618618618618p-5;   /* 1/21 */
    t = t * s2 + 0x1.af286bca1af28p-5;   /* 1/19 */
    t = t * s2 + 0x1.e1e1e1e1e1e1ep-5;   /* 1/17 */
    t = t * s2 + 0x1.1111111111111p-4;   /* 1/15 */
    t = t * s2 + 0x1.3b13b13b13b14p-4;   /* 1/13 */
    t = t * s2 + 0x1.745d1745d1746p-4;   /* 1/11 */
    t = t * s2 + 0x1.c71c71c71c71cp-4;   /* 1/9 */
    t = t * s2 + 0x1.2492492492492p-3;   /* 1/7 */
    t = t * s2 + 0x1.999999999999ap-3;   /* 1/5 */
    t = t * s2 + 0x1.5555555555555p-2;   /* 1/3 */
    t = t * s2;
    /* log(x) = e*ln2 + 2s + 2slo + 2s*t as lh + ll */
    two_sum((double)e * ln2_hi, 2.0 * s, &lh, &ll);
    ll = ll + ((double)e * ln2_lo + (2.0 * slo + 2.0 * s * t));
    tmp = lh + ll;
    ll = ll - (tmp - lh);
    lh = tmp;

    /* y*log(x) = th + tl */
    two_prod(y, lh, &th, &tl);
    tl = tl + y * ll;
    tmp = th + tl;
    tl = tl - (tmp - th);
    th = tmp;

    if (th > 709.0) return INFINITY;
    if (th < -745.0) return 0.0;

    /* exp(th + tl) = 2^k * exp(r) */
    k = (int)floor(th * inv_ln2 + 0.5);
    rh = th - (double)k * ln2_hi;
    r = (rh - (double)k * ln2_lo) + tl;
    q = 0x1.6124613a86d09p-33;          /* 1/13! */
    q = q * r + 0x1.1eed8eff8d898p-29;  /* 1/12! */
    q = q * r + 0x1.ae64567f544e4p-26;  /* 1/11! */
    q = q * r + 0x1.27e4fb7789f5cp-22;  /* 1/10! */
    q = q * r + 0x1.71de3a556c734p-19;  /* 1/9! */
    q = q * r + 0x1.a01a01a01a01ap-16;  /* 1/8! */
    q = q * r + 0x1.a01a01a01a01ap-13;  /* 1/7! */
    q = q * r + 0x1.6c16c16c16c17p-10;  /* 1/6! */
    q = q * r + 0x1.1111111111111p-7;   /* 1/5! */
    q = q * r + 0x1.5555555555555p-5;   /* 1/4! */
    q = q * r + 0x1.5555555555555p-3;   /* 1/3! */
    q = q * r + 0.5;                    /* 1/2! */
    ex = 1.0 + (r + (r * r) * q);
    /* scale by 2^k in two steps so that k in [-1100, 1100] stays exact */
    {
        int k1 = k / 2, k2 = k - k1;
        or_bits s1, s2b;
        s1.u = (uint64_t)(k1 + 1023) << 52;
        s2b.u = (uint64_t)(k2 + 1023) << 52;
        scale = s1.d;
        ex = ex * scale;
        ex = ex * s2b.d;
    }
    return ex;
}

/* ------------------------------------------------------------------------ */
/* Mesh preprocessing                                                        */

/* set_boundary + find_neighbour (maillage.cuf:850-914): the neighbour across
 * face k (nodes k -> k+1) of cell gi is the cell gj that holds the reversed
 * directed edge; otherwise the cell itself.  Column 2 starts as the cell id. */
void or_neighbours_allpairs(int nelt, const int *conn, int *b1, int *b2)
{
    int gi, gj;
    for (gi = 0; gi < nelt; gi++)
        for (int k = 0; k < 3; k++) {
            b1[3 * gi + k] = gi + 1;
            b2[3 * gi + k] = gi + 1;
        }
    for (gi = 0; gi < nelt; gi++) {
        int n1 = conn[3 * gi], n2 = conn[3 * gi + 1], n3 = conn[3 * gi + 2];
        for (gj = 0; gj < nelt; gj++) {
            int m1, m2, m3;
            if (gj == gi) continue;
            m1 = conn[3 * gj]; m2 = conn[3 * gj + 1]; m3 = conn[3 * gj + 2];
            if ((m1 == n1 && m3 == n2) || (m3 == n1 && m2 == n2) || (m2 == n1 && m1 == n2))
                b1[3 * gi + 0] = gj + 1;
            if ((m1 == n2 && m3 == n3) || (m3 == n2 && m2 == n3) || (m2 == n2 && m1 == n3))
                b1[3 * gi + 1] = gj + 1;
            if ((m1 == n3 && m3 == n1) || (m3 == n3 && m2 == n1) || (m2 == n3 && m1 == n1))
                b1[3 * gi + 2] = gj + 1;
        }
    }
}

typedef struct { int64_t key; int cell; } or_edge;

static int or_edge_cmp(const void *a, const void *b)
{
    const or_edge *x = (const or_edge *)a, *y = (const or_edge *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->cell - y->cell;
}

/* The same rule as or_neighbours_allpairs via a sorted directed-edge list. */
void or_neighbours_sorted(int nelt, const int *conn, int *b1, int *b2)
{
    size_t ne = (size_t)nelt * 3, i;
    or_edge *ed = (or_edge *)malloc(ne * sizeof(or_edge));
    for (int gi = 0; gi < nelt; gi++)
        for (int k = 0; k < 3; k++) {
            int a = conn[3 * gi + k], b = conn[3 * gi + (k + 1) % 3];
            ed[3 * gi + k].key = ((int64_t)a << 32) | (uint32_t)b;
            ed[3 * gi + k].cell = gi;
            b1[3 * gi + k] = gi + 1;
            b2[3 * gi + k] = gi + 1;
        }
    qsort(ed, ne, sizeof(or_edge), or_edge_cmp);
    for (int gi = 0; gi < nelt; gi++)
        for (int k = 0; k < 3; k++) {
            int a = conn[3 * gi + k], b = conn[3 * gi + (k + 1) % 3];
            int64_t want = ((int64_t)b << 32) | (uint32_t)a;
            size_t lo = 0, hi = ne;
            while (lo < hi) {
                size_t mid = (lo + hi) / 2;
                if (ed[mid].key < want) lo = mid + 1; else hi = mid;
            }
            /* highest matching cell other than gi */
            for (i = lo; i < ne && ed[i].key == want; i++)
                if (ed[i].cell != gi) b1[3 * gi + k] = ed[i].cell + 1;
        }
    free(ed);
}

static int *or_positions(int nmax, int nd, const int *list)
{
    int *pos = (int *)malloc((size_t)(nmax + 1) * sizeof(int));
    for (int i = 0; i <= nmax; i++) pos[i] = -1;
    for (int i = nd - 1; i >= 0; i--)
        if (list[i] >= 0 && list[i] <= nmax) pos[list[i]] = i;
    return pos;
}

/* boundary_kind (maillage.cuf:981-1110) with assign_type_1/_numero
 * (:1123-1173) and assign_type_2 (:1201-1225): a face whose neighbour is the
 * cell itself is an inlet (-1) if both nodes are inlet nodes, else an outlet
 * (-2) if both are outlet nodes, else a wall (-3) if both are wall nodes, else
 * a wall if one node is a wall node and the other an inlet or outlet node. */
void or_boundary_kind(int nelt, const int *conn, int *b1, int *b2,
                      int ndi, const int *ndinput, const int *num_in,
                      int ndo, const int *ndoutput, const int *num_out,
                      int ndw, const int *ndwall)
{
    int nmax = 0;
    for (int i = 0; i < 3 * nelt; i++) if (conn[i] > nmax) nmax = conn[i];
    int *pi = or_positions(nmax, ndi, ndinput);
    int *po = or_positions(nmax, ndo, ndoutput);
    int *pw = or_positions(nmax, ndw, ndwall);
    for (int kvol = 0; kvol < nelt; kvol++)
        for (int kfac = 0; kfac < 3; kfac++) {
            int n1, n2, a, c;
            if (b1[3 * kvol + kfac] != kvol + 1) continue;
            n1 = conn[3 * kvol + kfac];
            n2 = conn[3 * kvol + (kfac + 1) % 3];
            if (ndi > 0) {
                a = pi[n1]; c = pi[n2];
                if (n1 != n2 && a >= 0 && c >= 0) {
                    b1[3 * kvol + kfac] = -1;
                    b2[3 * kvol + kfac] = num_in[a < c ? a : c];
                    continue;
                }
            }
            if (ndo > 0) {
                a = po[n1]; c = po[n2];
                if (n1 != n2 && a >= 0 && c >= 0) {
                    b1[3 * kvol + kfac] = -2;
                    b2[3 * kvol + kfac] = num_out[a < c ? a : c];
                    continue;
                }
            }
            if (ndw > 0) {
                a = pw[n1]; c = pw[n2];
                if (n1 != n2 && a >= 0 && c >= 0) {
                    b1[3 * kvol + kfac] = -3;
                    continue;
                }
            }
            if (ndi > 0 && ndw > 0)
                if ((pw[n1] >= 0 && pi[n2] >= 0) || (pi[n1] >= 0 && pw[n2] >= 0))
                    b1[3 * kvol + kfac] = -3;
            if (ndo > 0 && ndw > 0)
                if ((pw[n1] >= 0 && po[n2] >= 0) || (po[n1] >= 0 && pw[n2] >= 0))
                    b1[3 * kvol + kfac] = -3;
        }
    free(pi); free(po); free(pw);
}

/* set_undefined_boudaries_as_walls (maillage.cuf:1454-1471) */
void or_set_walls(int nelt, int *b1)
{
    for (int gi = 0; gi < nelt; gi++)
        for (int k = 0; k < 3; k++)
            if (b1[3 * gi + k] == gi + 1) b1[3 * gi + k] = -3;
}

/* find_upwind_cells (maillage.cuf:954-978): for face c the two other
 * neighbours (faces c+1, c+2), or the cell itself across a boundary. */
void or_upwind_cells(int nelt, const int *b1, int *upw)
{
    for (int gi = 0; gi < nelt; gi++)
        for (int c = 0; c < 3; c++) {
            int f1 = b1[3 * gi + (c + 1) % 3], f2 = b1[3 * gi + (c + 2) % 3];
            upw[6 * gi + 2 * c + 0] = f1 > 0 ? f1 : gi + 1;
            upw[6 * gi + 2 * c + 1] = f2 > 0 ? f2 : gi + 1;
        }
}

/* construct_coo_table_elemwise (maillage.cuf:1254-1282), the cotemin loop
 * (:132-159) and compute_surface_and_bathy (:1284-1331). */
void or_geometry(int nelt, const int *conn, const double *x, const double *y,
                 const double *z, double *vx, double *vy, double *xc, double *yc,
                 double *surf, double *zm, double *cotemin_arr, double *cotemin)
{
    double cmin = 1000000.0;
    for (int i = 0; i < nelt; i++) {
        int s1 = conn[3 * i] - 1, s2 = conn[3 * i + 1] - 1, s3 = conn[3 * i + 2] - 1;
        double x1 = x[s1], y1 = y[s1], z1 = z[s1];
        double x2 = x[s2], y2 = y[s2], z2 = z[s2];
        double x3 = x[s3], y3 = y[s3], z3 = z[s3];
        double a, b, c, s, ca, cb, cc, p, aire, xg, yg, d1, d2, d3, cot;
        vx[3 * i] = x1; vx[3 * i + 1] = x2; vx[3 * i + 2] = x3;
        vy[3 * i] = y1; vy[3 * i + 1] = y2; vy[3 * i + 2] = y3;
        xc[i] = ((x1 + x2) + x3) / 3.0;
        yc[i] = ((y1 + y2) + y3) / 3.0;
        /* cotemin_arr = 1.8 * inscribed radius (maillage.cuf:148-157) */
        ca = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
        cb = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3));
        cc = sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1));
        p = ((ca + cb) + cc) / 2.0;
        aire = sqrt(((p * (p - ca)) * (p - cb)) * (p - cc));
        cot = (0.5 * aire) / p;
        cot = LIT_1P8 * cot;
        cotemin_arr[i] = cot;
        cmin = dmin(cmin, cot);
        /* Heron area and distance-weighted bed (maillage.cuf:1307-1329) */
        a = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
        b = sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
        c = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3));
        s = ((a + b) + c) / 2.0;
        surf[i] = sqrt(((s * (s - a)) * (s - b)) * (s - c));
        xg = ((x1 + x2) + x3) / 3.0;
        yg = ((y1 + y2) + y3) / 3.0;
        d1 = sqrt((x1 - xg) * (x1 - xg) + (y1 - yg) * (y1 - yg));
        d2 = sqrt((x2 - xg) * (x2 - xg) + (y2 - yg) * (y2 - yg));
        d3 = sqrt((x3 - xg) * (x3 - xg) + (y3 - yg) * (y3 - yg));
        zm[i] = ((d1 * z1 + d2 * z2) + d3 * z3) / ((d1 + d2) + d3);
    }
    if (cotemin) *cotemin = cmin;
}

/* calcul_long_entree (maillage.cuf:1423-1452) */
void or_long_entree(int nelt, const int *conn, const double *x, const double *y,
                    const int *b1, const int *b2, int nentrees, double *long_entree)
{
    for (int k = 0; k < nentrees; k++) long_entree[k] = 0.0;
    for (int kvol = 0; kvol < nelt; kvol++)
        for (int kfac = 0; kfac < 3; kfac++) {
            int n1 = conn[3 * kvol + kfac] - 1, n2 = conn[3 * kvol + (kfac + 1) % 3] - 1;
            if (b1[3 * kvol + kfac] == -1) {
                int id = b2[3 * kvol + kfac];
                if (id >= 1 && id <= nentrees)
                    long_entree[id - 1] = long_entree[id - 1] +
                        sqrt((x[n1] - x[n2]) * (x[n1] - x[n2]) + (y[n1] - y[n2]) * (y[n1] - y[n2]));
            }
        }
}

/* set_initial_condition, dam-line branch (initial_condition.cuf:31-64) */
void or_ic_dam(int nelt, const double *xc, const double *yc, const double *zm,
               double tolisec, double x1, double y1, double x2, double y2,
               double h_amont, double u_amont, double v_amont,
               double h_aval, double u_aval, double v_aval, double *U)
{
    double alpha = (y1 - y2) / (x1 - x2);
    double beta = y1 - alpha * x1;
    for (int gi = 0; gi < nelt; gi++) {
        double h, u, v;
        if ((alpha * xc[gi] - yc[gi]) + beta >= 0.0) { h = h_amont - zm[gi]; u = u_amont; v = v_amont; }
        else { h = h_aval - zm[gi]; u = u_aval; v = v_aval; }
        if (h > tolisec) { U[3 * gi] = h; U[3 * gi + 1] = h * u; U[3 * gi + 2] = h * v; }
        else { U[3 * gi] = 0.0; U[3 * gi + 1] = 0.0; U[3 * gi + 2] = 0.0; }
    }
}

/* ------------------------------------------------------------------------ */
/* Riemann solvers                                                           */

/* HLL wave speeds and the HLLC flux in the face frame, shared by the HLLC
 * solver (flux_riemann.cuf:930-1020) and the first half of WAF (:753-845).
 * Returns 0 for a dry/dry face (zero flux, see header). */
static int or_hllc_core(double gp, double eps, double hl, double ul, double vl,
                        double hr, double ur, double vr, double f1[3], double f2[3],
                        double fh[3], double *psl, double *psr, double *pset)
{
    double al, ar, het, uet, pql, pqr, sl, sr, set, denom, q1[3], q2[3];
    if (hl < eps && hr < eps) return 0;
    al = sqrt(gp * hl);
    ar = sqrt(gp * hr);
    het = (hl + hr) / 2.0 - (((ur - ul) * (hl + hr)) / (al + ar)) / 4.0;
    uet = (ul + ur) / 2.0 - ((hr - hl) * (al + ar)) / (hl + hr);
    if (het < hl) pql = 1.0;
    else pql = (hl > eps) ? sqrt(((0.5 * (het + hl)) * het) / (hl * hl)) : 0.0;
    if (het < hr) pqr = 1.0;
    else pqr = (hr > eps) ? sqrt(((0.5 * (het + hr)) * het) / (hr * hr)) : 0.0;
    if (hl > eps) {
        sl = ul - al * pql;
        if (hr > eps) sr = ur + ar * pqr;
        else { sl = ul - al; sr = ul + 2.0 * al; }
    } else {
        sl = ur - 2.0 * ar;
        sr = ur + ar;
    }
    denom = hr * (ur - sr) - hl * (ul - sl);
    if (fabs(denom) < eps) set = uet;
    else set = ((sl * hr) * (ur - sr) - (sr * hl) * (ul - sl)) / denom;
    q1[0] = hl; q1[1] = hl * ul; q1[2] = hl * vl;
    q2[0] = hr; q2[1] = hr * ur; q2[2] = hr * vr;
    f1[0] = hl * ul; f1[1] = hl * (ul * ul) + ((gp * hl) * hl) / 2.0; f1[2] = (hl * ul) * vl;
    f2[0] = hr * ur; f2[1] = hr * (ur * ur) + ((gp * hr) * hr) / 2.0; f2[2] = (hr * ur) * vr;
    if (sl > 0.0) { fh[0] = f1[0]; fh[1] = f1[1]; fh[2] = f1[2]; }
    else if (sr < 0.0) { fh[0] = f2[0]; fh[1] = f2[1]; fh[2] = f2[2]; }
    else {
        for (int k = 0; k < 2; k++)
            fh[k] = ((sr * f1[k] - sl * f2[k]) + (sl * sr) * (q2[k] - q1[k])) / ((sr - sl) + eps);
        fh[2] = (set > 0.0) ? fh[0] * vl : fh[0] * vr;
    }
    *psl = sl; *psr = sr; *pset = set;
    return 1;
}

/* solver_Riemann_hllc (flux_riemann.cuf:902-1026) */
static void or_hllc(double gp, double n1, double n2, double hl, double ul, double vl,
                    double hr, double ur, double vr, double flux[3])
{
    double f1[3], f2[3], fh[3] = {0.0, 0.0, 0.0}, sl, sr, set;
    or_hllc_core(gp, LIT_1EM8, hl, ul, vl, hr, ur, vr, f1, f2, fh, &sl, &sr, &set);
    flux[0] = fh[0];
    flux[1] = n1 * fh[1] - n2 * fh[2];
    flux[2] = n2 * fh[1] + n1 * fh[2];
}

/* solver_Riemann_riadh (flux_riemann.cuf:1029-1182), iflux 2: the same
 * wave speeds as the HLLC solver with eps = 1e-6f, then the flux of the
 * region holding x/t = 0 among U_L, U*_L, U*_R, U_R, the star fluxes being
 * F_K + S_K (U*_K - U_K) with U*_K = h_K (S_K - u_K)/(S_K - S*) (1, S*, v_K)
 * (the weight set to 0 when |S_K - S*| <= eps). */
static void or_riadh(double gp, double n1, double n2, double hl, double ul, double vl,
                     double hr, double ur, double vr, double flux[3])
{
    const double eps = LIT_1EM6;
    double al, ar, het, uet, pql, pqr, sl, sr, set, denom, pond;
    double q1[3], q2[3], f1[3], f2[3], e1[3], e2[3], fx[3];
    if (hl < eps && hr < eps) {  /* both dry (:1055): zero flux, see header */
        flux[0] = flux[1] = flux[2] = 0.0;
        return;
    }
    al = sqrt(gp * hl);
    ar = sqrt(gp * hr);
    het = (hl + hr) / 2.0 - (((ur - ul) * (hl + hr)) / (al + ar)) / 4.0;
    uet = (ul + ur) / 2.0 - ((hr - hl) * (al + ar)) / (hl + hr);
    if (het < hl) pql = 1.0;
    else pql = (hl > eps) ? sqrt(((0.5 * (het + hl)) * het) / (hl * hl)) : 0.0;
    if (het < hr) pqr = 1.0;
    else pqr = (hr > eps) ? sqrt(((0.5 * (het + hr)) * het) / (hr * hr)) : 0.0;
    if (hl > eps) {
        sl = ul - al * pql;
        if (hr > eps) sr = ur + ar * pqr;
        else { sl = ul - al; sr = ul + 2.0 * al; }
    } else {
        sl = ur - 2.0 * ar;
        sr = ur + ar;
    }
    denom = hr * (ur - sr) - hl * (ul - sl);
    if (fabs(denom) < eps) set = uet;
    else set = ((sl * hr) * (ur - sr) - (sr * hl) * (ul - sl)) / denom;
    q1[0] = hl; q1[1] = hl * ul; q1[2] = hl * vl;
    q2[0] = hr; q2[1] = hr * ur; q2[2] = hr * vr;
    pond = (fabs(sl - set) > eps) ? (hl * (sl - ul)) / (sl - set) : 0.0;
    e1[0] = pond; e1[1] = pond * set; e1[2] = pond * vl;
    pond = (fabs(sr - set) > eps) ? (hr * (sr - ur)) / (sr - set) : 0.0;
    e2[0] = pond; e2[1] = pond * set; e2[2] = pond * vr;
    f1[0] = hl * ul; f1[1] = hl * (ul * ul) + ((gp * hl) * hl) / 2.0; f1[2] = (hl * ul) * vl;
    f2[0] = hr * ur; f2[1] = hr * (ur * ur) + ((gp * hr) * hr) / 2.0; f2[2] = (hr * ur) * vr;
    for (int k = 0; k < 3; k++) {
        e1[k] = f1[k] + sl * (e1[k] - q1[k]);
        e2[k] = f2[k] + sr * (e2[k] - q2[k]);
    }
    if (0.0 < sl) { fx[0] = f1[0]; fx[1] = f1[1]; fx[2] = f1[2]; }
    else if (0.0 < set && 0.0 > sl) { fx[0] = e1[0]; fx[1] = e1[1]; fx[2] = e1[2]; }
    else if (0.0 > set && 0.0 < sr) { fx[0] = e2[0]; fx[1] = e2[1]; fx[2] = e2[2]; }
    else { fx[0] = f2[0]; fx[1] = f2[1]; fx[2] = f2[2]; }
    flux[0] = fx[0];
    flux[1] = n1 * fx[1] - n2 * fx[2];
    flux[2] = n2 * fx[1] + n1 * fx[2];
}

/* limiter (flux_riemann.cuf:465-524) */
static double or_limiter(int ilim, double r, double c)
{
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

/* solver_Riemann_waf (flux_riemann.cuf:720-900) */
static void or_waf(const or_params *p, double n1, double n2,
                   double hl, double ul, double vl, double hr, double ur, double vr,
                   double hl_upw, double vl_upw, double hr_upw, double vr_upw,
                   double d_mrl, double d_mkvol, double d_mkvoisin, double dt,
                   double flux[3])
{
    const double eps = LIT_1EM16;
    double f1[3], f2[3], fh[3], sl, sr, set, fw[3];
    double dtoverdx, cr, cl, cet, dlr, dll, dlet, dur, dul, duet, rl, rr, ret;
    double lim_l, lim_r, lim_et, wl, wlet, wlr, wr, wret;
    if (!or_hllc_core(p->gp, eps, hl, ul, vl, hr, ur, vr, f1, f2, fh, &sl, &sr, &set)) {
        fw[0] = 0.0; fw[1] = 0.0; fw[2] = 0.0;
    } else {
        dtoverdx = dt / d_mrl;
        cr = sr * dtoverdx;
        cl = sl * dtoverdx;
        cet = set * dtoverdx;
        dlr = (hr - hl) / d_mrl;
        dll = (hr - hl) / d_mrl;
        dlet = (vr - vl) / d_mrl;
        dur = (sr > 0.0) ? (hl - hl_upw) / d_mkvol : (hr_upw - hr) / d_mkvoisin;
        dul = (sl > 0.0) ? (hl - hl_upw) / d_mkvol : (hr_upw - hr) / d_mkvoisin;
        duet = (set > 0.0) ? (vl - vl_upw) / d_mkvol : (vr_upw - vr) / d_mkvoisin;
        rl = dul / (dll + eps);
        rr = dur / (dlr + eps);
        ret = duet / (dlet + eps);
        lim_l = or_limiter(p->ilimiteur, rl, cl);
        lim_r = or_limiter(p->ilimiteur, rr, cr);
        lim_et = or_limiter(p->ilimiteur, ret, cet);
        wl = (1.0 + copysign(1.0, sl) * lim_l) / 2.0;
        wlet = (1.0 + copysign(1.0, set) * lim_et) / 2.0;
        wlr = (copysign(1.0, sr) * lim_r - copysign(1.0, sl) * lim_l) / 2.0;
        wr = (1.0 - copysign(1.0, sr) * lim_r) / 2.0;
        wret = (1.0 - copysign(1.0, set) * lim_et) / 2.0;
        fw[0] = (wl * f1[0] + wlr * fh[0]) + wr * f2[0];
        fw[1] = (wl * f1[1] + wlr * fh[1]) + wr * f2[1];
        fw[2] = (wlet * vl + wret * vr) * fw[0];
    }
    flux[0] = fw[0];
    flux[1] = n1 * fw[1] - n2 * fw[2];
    flux[2] = n2 * fw[1] + n1 * fw[2];
}

/* simple_limiter (flux_riemann.cuf:430-462): the MUSCL slope from the
 * upwind slope r1 and the slope across the face r2.  The reference leaves
 * the slope unset for ilimiteur outside 0..3: 0 here. */
static double or_slope_limiter(int ilim, double r1, double r2)
{
    switch (ilim) {
    case 1:  /* superbee */
        if (r2 > 0.0) return dmax(dmax(0.0, dmin(2.0 * r1, r2)), dmin(r1, 2.0 * r2));
        return dmin(dmin(0.0, dmax(2.0 * r1, r2)), dmax(r1, 2.0 * r2));
    case 2:  /* van Leer */
        return (r1 * r2 + fabs(r1 * r2)) / ((r1 + r2) + LIT_1EM12);
    case 3:  /* minmod */
        if (r2 > 0.0) return dmax(0.0, dmin(r1, r2));
        return dmin(0.0, dmax(r1, r2));
    default:
        return 0.0;
    }
}

/* solver_Riemann_muscl (flux_riemann.cuf:526-718), iflux 4: limited linear
 * reconstruction of each side to its two face values (:560-610), half-step
 * evolution with the difference of the face fluxes (:613-620, velocities
 * evolved with momentum-flux differences as written there), then the HLLC
 * solver with eps = 1e-16f on the evolved inner values (:622-717). */
static void or_muscl(const or_params *p, double n1, double n2,
                     double hl, double ul, double vl, double hr, double ur, double vr,
                     double hl_upw, double ul_upw, double vl_upw,
                     double hr_upw, double ur_upw, double vr_upw,
                     double d_mrl, double d_mkvol, double d_mkvoisin, double dt, double flux[3])
{
    const double tls = p->tolisec, gp = p->gp;
    const int il = p->ilimiteur;
    double dtoverdx = dt / d_mrl;
    double d, hll, hlr, ull, ulr, vll, vlr, hrl, hrr, url, urr, vrl, vrr;
    double eh, eu, ev, fhh, fu, fv, f1[3], f2[3], fh[3] = {0.0, 0.0, 0.0}, sl, sr, set;

    d = or_slope_limiter(il, (hl - hl_upw) / d_mkvol, (hr - hl) / d_mrl);
    hll = dmax(0.0, hl - (0.5 * d) * d_mrl);
    hlr = dmax(0.0, hl + (0.5 * d) * d_mrl);
    if (hlr < tls || hll < tls) {
        ulr = 0.0; vlr = 0.0; ull = 0.0; vll = 0.0;
    } else {
        double du = or_slope_limiter(il, (ul - ul_upw) / d_mkvol, (ur - ul) / d_mrl);
        double dv = or_slope_limiter(il, (vl - vl_upw) / d_mkvol, (vr - vl) / d_mrl);
        ull = ul - (0.5 * du) * d_mrl;
        vll = vl - (0.5 * dv) * d_mrl;
        ulr = ul + (0.5 * du) * d_mrl;
        vlr = vl + (0.5 * dv) * d_mrl;
    }
    d = or_slope_limiter(il, (hr - hl) / d_mrl, (hr_upw - hr) / d_mkvoisin);
    hrr = dmax(0.0, hr + (0.5 * d) * d_mrl);
    hrl = dmax(0.0, hr - (0.5 * d) * d_mrl);
    if (hrl < tls || hrr < tls) {
        url = 0.0; vrl = 0.0; urr = 0.0; vrr = 0.0;
    } else {
        double du = or_slope_limiter(il, (ur - ul) / d_mrl, (ur_upw - ur) / d_mkvoisin);
        double dv = or_slope_limiter(il, (vr - vl) / d_mrl, (vr_upw - vr) / d_mkvoisin);
        urr = ur + (0.5 * du) * d_mrl;
        vrr = vr + (0.5 * dv) * d_mrl;
        url = ur - (0.5 * du) * d_mrl;
        vrl = vr - (0.5 * dv) * d_mrl;
    }
    eh = hlr + (0.5 * dtoverdx) * (hll * ull - hlr * ulr);
    eu = ulr + (0.5 * dtoverdx) * ((hll * (ull * ull) + (0.5 * gp) * (hll * hll)) -
                                   (hlr * (ulr * ulr) + (0.5 * gp) * (hlr * hlr)));
    ev = vlr + (0.5 * dtoverdx) * ((hll * ull) * vll - (hlr * ulr) * vlr);
    fhh = hrl + (0.5 * dtoverdx) * (hrl * url - hrr * urr);
    fu = url + (0.5 * dtoverdx) * ((hrl * (url * url) + (0.5 * gp) * (hrl * hrl)) -
                                   (hrr * (urr * urr) + (0.5 * gp) * (hrr * hrr)));
    fv = vrl + (0.5 * dtoverdx) * ((hrl * url) * vrl - (hrr * urr) * vrr);
    or_hllc_core(gp, LIT_1EM16, eh, eu, ev, fhh, fu, fv, f1, f2, fh, &sl, &sr, &set);
    flux[0] = fh[0];
    flux[1] = n1 * fh[1] - n2 * fh[2];
    flux[2] = n2 * fh[1] + n1 * fh[2];
}

/* ------------------------------------------------------------------------ */
/* Per-step kernels                                                          */

/* Inflow discharge at time tc.  The reference's debitglob is constant
 * (data_transfer.cuf:76); a hydrograph table (t_i, Q_i), linearly
 * interpolated and held constant outside its range, is this build's
 * extension for time-varying inflow. */
double or_inflow_q(const or_params *p, int id, double tc)
{
    int n = p->hydro_n[id];
    const double *t = p->hydro_t + id * OR_MAX_HYDRO, *q = p->hydro_q + id * OR_MAX_HYDRO;
    if (n <= 0) return p->debitglob[id];
    if (tc <= t[0]) return q[0];
    for (int i = 0; i + 1 < n; i++)
        if (tc < t[i + 1]) return q[i] + (q[i + 1] - q[i]) * ((tc - t[i]) / (t[i + 1] - t[i]));
    return q[n - 1];
}

/* Rotated (normal, tangential) velocity of a cell with the dry threshold of
 * flux_riemann.cuf:105-112 / :186-193 / :240-258. */
static inline void or_rot(double tolisec, double n1, double n2, const double *u,
                          double *h, double *un, double *ut)
{
    *h = u[0];
    if (*h > tolisec) {
        *un = (u[1] * n1 + u[2] * n2) / *h;
        *ut = (u[2] * n1 - u[1] * n2) / *h;
    } else { *h = 0.0; *un = 0.0; *ut = 0.0; }
}

static inline double or_cdist(const double *xc, const double *yc, int a, int b)
{
    return sqrt((xc[a] - xc[b]) * (xc[a] - xc[b]) + (yc[a] - yc[b]) * (yc[a] - yc[b]));
}

/* Outward unit normal and length of face tj (flux_riemann.cuf:50-74,
 * source_terms.cuf:62-88). */
static inline void or_face_geom(const double *vx, const double *vy, int gi, int tj,
                                double *n1, double *n2, double *ln)
{
    int c1 = tj, c2 = (tj + 1) % 3, c3 = (tj + 2) % 3;
    double x1 = vx[3 * gi + c1], y1 = vy[3 * gi + c1];
    double x2 = vx[3 * gi + c2], y2 = vy[3 * gi + c2];
    double x3 = vx[3 * gi + c3], y3 = vy[3 * gi + c3];
    double l = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    double a = (y2 - y1) / l, b = (x1 - x2) / l;
    double orient = a * ((x1 + x2) / 2.0 - ((x1 + x2) + x3) / 3.0) +
                    b * ((y1 + y2) / 2.0 - ((y1 + y2) + y3) / 3.0);
    if (orient < 0.0) { a = (y1 - y2) / l; b = (x2 - x1) / l; }
    *n1 = a; *n2 = b; *ln = l;
}

/* compute_flux (flux_riemann.cuf:4-429) for one (cell, face): returns
 * flux*ln components; *qo receives the outflow edge discharge
 * (debit_sortie_arr, flux_riemann.cuf:161) and *qi the inflow edge
 * discharge (debit_entree_arr, :125-126). */
static void or_face(const or_params *p, const int *conn, const double *vx,
                    const double *vy, const double *xc, const double *yc,
                    const double *zm, const int *b1, const int *b2,
                    const int *upw, const double *U0, double dt, double tc,
                    int gi, int tj, double fl[3], double *qo, double *qi)
{
    const double g = p->gp, tls = p->tolisec;
    double n1, n2, ln, flux[3] = {0.0, 0.0, 0.0};
    double hautl, vitul, vitvl, zml;
    int kvoisin = b1[3 * gi + tj];
    int id = b2[3 * gi + tj];
    or_face_geom(vx, vy, gi, tj, &n1, &n2, &ln);
    zml = zm[gi];
    or_rot(tls, n1, n2, U0 + 3 * gi, &hautl, &vitul, &vitvl);
    *qo = 0.0;
    *qi = 0.0;

    if (kvoisin == -1) {
        if (p->inlet == 1) {
            flux[0] = hautl * vitul;
            flux[1] = n1 * (hautl * (vitul * vitul) + (g * (hautl * hautl)) / 2.0) - ((n2 * hautl) * vitul) * vitvl;
            flux[2] = n2 * (hautl * (vitul * vitul) + (g * (hautl * hautl)) / 2.0) + ((n1 * hautl) * vitul) * vitvl;
        } else if (p->inlet == 2) {
            double dn = -(or_inflow_q(p, id - 1, tc) / p->long_entree[id - 1]);
            flux[0] = dn;
            flux[1] = n1 * ((dn * dn) / (hautl + tls) + (g * (hautl * hautl)) / 2.0);
            flux[2] = n2 * ((dn * dn) / (hautl + tls) + (g * (hautl * hautl)) / 2.0);
            *qi = (-dn) * ln;
        } else if (p->inlet == 3) {
            /* phys_inflow (flux_riemann.cuf:127-137): the inlet discharge
               shared out in proportion to the depth */
            double dn = 0.0;
            if (p->surf_entree[id - 1] > LIT_1EM8)
                dn = -((or_inflow_q(p, id - 1, tc) * hautl) / p->surf_entree[id - 1]);
            flux[0] = dn;
            flux[1] = n1 * ((dn * dn) / (hautl + tls) + (g * (hautl * hautl)) / 2.0);
            flux[2] = n2 * ((dn * dn) / (hautl + tls) + (g * (hautl * hautl)) / 2.0);
            *qi = (-dn) * ln;
        }
    }
    if (kvoisin == -3) {
        flux[0] = 0.0;
        flux[1] = ((n1 * g) * (hautl * hautl)) / 2.0;
        flux[2] = ((n2 * g) * (hautl * hautl)) / 2.0;
    }
    if (kvoisin == -2 && p->outlet == 1) {
        if (vitul >= 0.0 && vitul / sqrt(g * hautl) < 1.0) {
            hautl = p->h_sortie[id - 1] - zml;
            hautl = dmax(0.0, hautl);
            flux[0] = hautl * vitul;
            flux[1] = n1 * (hautl * (vitul * vitul) + (g * (hautl * hautl)) / 2.0) - ((n2 * hautl) * vitul) * vitvl;
            flux[2] = n2 * (hautl * (vitul * vitul) + (g * (hautl * hautl)) / 2.0) + ((n1 * hautl) * vitul) * vitvl;
        } else {
            flux[0] = 0.0;
            flux[1] = ((n1 * g) * (hautl * hautl)) / 2.0;
            flux[2] = ((n2 * g) * (hautl * hautl)) / 2.0;
        }
        *qo = flux[0] * ln;
    }
    if (kvoisin > 0 || (kvoisin == -2 && p->outlet == 2)) {
        int kv = (kvoisin > 0) ? kvoisin - 1 : gi;
        double hautr, vitur, vitvr, zmr, zij, hijet, hjiet, d_mrl;
        zmr = zm[kv];
        zij = dmax(zml, zmr);
        hijet = dmax(0.0, (hautl + zml) - zij);
        hjiet = dmax(0.0, (U0[3 * kv] + zmr) - zij);
        or_rot(tls, n1, n2, U0 + 3 * kv, &hautr, &vitur, &vitvr);
        d_mrl = or_cdist(xc, yc, gi, kv);
        if (p->iflux == 1) {
            if (hijet < tls) { hijet = 0.0; vitul = 0.0; vitvl = 0.0; }
            if (hjiet < tls) { hjiet = 0.0; vitur = 0.0; vitvr = 0.0; }
            or_hllc(g, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr, flux);
        } else if (p->iflux == 2) {  /* flux_riemann.cuf:213-228 */
            if (hijet < tls) { hijet = 0.0; vitul = 0.0; vitvl = 0.0; }
            if (hjiet < tls) { hjiet = 0.0; vitur = 0.0; vitvr = 0.0; }
            or_riadh(g, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr, flux);
        } else if (p->iflux == 3 || p->iflux == 4) {  /* shared stencil (:229-394) */
            int kv1 = upw[6 * gi + 2 * tj] - 1, kv2 = upw[6 * gi + 2 * tj + 1] - 1;
            int A = conn[3 * gi + tj], B = conn[3 * gi + (tj + 1) % 3];
            int c1 = conn[3 * kv], c2 = conn[3 * kv + 1], c3 = conn[3 * kv + 2];
            int tjv, kw1, kw2;
            double h1, u1, w1, h2, u2, w2, g1h, g1u, g1w, g2h, g2u, g2w;
            double hl_upw, ul_upw, vl_upw, hr_upw, ur_upw, vr_upw, zml_upw, zmr_upw;
            double d_mkvol, d_mkvoisin;
            or_rot(tls, n1, n2, U0 + 3 * kv1, &h1, &u1, &w1);
            or_rot(tls, n1, n2, U0 + 3 * kv2, &h2, &u2, &w2);
            if ((c1 == A && c2 == B) || (c2 == A && c1 == B)) tjv = 0;
            else if ((c2 == A && c3 == B) || (c3 == A && c2 == B)) tjv = 1;
            else if ((c3 == A && c1 == B) || (c1 == A && c3 == B)) tjv = 2;
            else tjv = 0;
            kw1 = upw[6 * kv + 2 * tjv] - 1;
            kw2 = upw[6 * kv + 2 * tjv + 1] - 1;
            or_rot(tls, n1, n2, U0 + 3 * kw1, &g1h, &g1u, &g1w);
            or_rot(tls, n1, n2, U0 + 3 * kw2, &g2h, &g2u, &g2w);
            if ((kv1 == gi && kv2 == gi) || (kw1 == kv && kw2 == kv)) {
                hl_upw = hautl; ul_upw = vitul; vl_upw = vitvl;
                hr_upw = hautr; ur_upw = vitur; vr_upw = vitvr;
                d_mkvol = d_mrl; d_mkvoisin = d_mrl;
                zml_upw = zml; zmr_upw = zmr;
            } else {
                if (kv1 == gi) {
                    hl_upw = h2; ul_upw = u2; vl_upw = w2;
                    d_mkvol = or_cdist(xc, yc, kv2, gi);
                    zml_upw = zm[kv2];
                } else if (kv2 == gi) {
                    hl_upw = h1; ul_upw = u1; vl_upw = w1;
                    d_mkvol = or_cdist(xc, yc, kv1, gi);
                    zml_upw = zm[kv1];
                } else {
                    double d1 = or_cdist(xc, yc, kv1, gi), d2 = or_cdist(xc, yc, kv2, gi);
                    hl_upw = (d1 * h1 + d2 * h2) / (d1 + d2);
                    ul_upw = (d1 * u1 + d2 * u2) / (d1 + d2);
                    vl_upw = (d1 * w1 + d2 * w2) / (d1 + d2);
                    zml_upw = (d1 * zm[kv1] + d2 * zm[kv2]) / (d1 + d2);
                    d_mkvol = (d1 + d2) / 2.0;
                }
                if (kw1 == kv) {
                    hr_upw = g2h; ur_upw = g2u; vr_upw = g2w;
                    d_mkvoisin = or_cdist(xc, yc, kw2, kv);
                    zmr_upw = zm[kw2];
                } else if (kw2 == kv) {
                    hr_upw = g1h; ur_upw = g1u; vr_upw = g1w;
                    d_mkvoisin = or_cdist(xc, yc, kw1, kv);
                    zmr_upw = zm[kw1];
                } else {
                    double d1 = or_cdist(xc, yc, kw1, kv), d2 = or_cdist(xc, yc, kw2, kv);
                    hr_upw = (d1 * g1h + d2 * g2h) / (d1 + d2);
                    ur_upw = (d1 * g1u + d2 * g2u) / (d1 + d2);
                    vr_upw = (d1 * g1w + d2 * g2w) / (d1 + d2);
                    zmr_upw = (d1 * zm[kw1] + d2 * zm[kw2]) / (d1 + d2);
                    d_mkvoisin = (d1 + d2) / 2.0;
                }
            }
            /* hydrostatic reconstruction (flux_riemann.cuf:370-400) */
            zij = dmax(zml, zmr);
            hijet = dmax(0.0, (hautl + zml) - zij);
            hl_upw = dmax(0.0, (hl_upw + zml_upw) - zij);
            hjiet = dmax(0.0, (hautr + zmr) - zij);
            hr_upw = dmax(0.0, (hr_upw + zmr_upw) - zij);
            if (hijet < tls) { hijet = 0.0; vitul = 0.0; vitvl = 0.0; }
            if (hl_upw < tls) { hl_upw = 0.0; ul_upw = 0.0; vl_upw = 0.0; }
            if (hjiet < tls) { hjiet = 0.0; vitur = 0.0; vitvr = 0.0; }
            if (hr_upw < tls) { hr_upw = 0.0; ur_upw = 0.0; vr_upw = 0.0; }
            if (p->iflux == 3)
                or_waf(p, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr,
                       hl_upw, vl_upw, hr_upw, vr_upw, d_mrl, d_mkvol, d_mkvoisin, dt, flux);
            else
                or_muscl(p, n1, n2, hijet, vitul, vitvl, hjiet, vitur, vitvr, hl_upw, ul_upw, vl_upw,
                         hr_upw, ur_upw, vr_upw, d_mrl, d_mkvol, d_mkvoisin, dt, flux);
        }
    }
    fl[0] = flux[0] * ln;
    fl[1] = flux[1] * ln;
    fl[2] = flux[2] * ln;
}

/* compute_flux for all cells: resc(k,i) = (F1 + F2) + F3 (flux_riemann.cuf:411-426).
 * qout_cell[i] / qin_cell[i] (optional) receive the cell's outflow / inflow
 * discharge, summed over its faces in face order. */
void or_compute_flux(const or_params *p, int nelt, const int *conn,
                     const double *vx, const double *vy, const double *xc,
                     const double *yc, const double *zm, const int *b1,
                     const int *b2, const int *upw, const double *U0,
                     double dt, double tc, double *resc, double *qout_cell, double *qin_cell)
{
    int gi;
#pragma omp parallel for schedule(static)
    for (gi = 0; gi < nelt; gi++) {
        double f[3][3], q[3], qi[3];
        for (int tj = 0; tj < 3; tj++)
            or_face(p, conn, vx, vy, xc, yc, zm, b1, b2, upw, U0, dt, tc, gi, tj, f[tj], &q[tj], &qi[tj]);
        for (int k = 0; k < 3; k++) resc[3 * gi + k] = (f[0][k] + f[1][k]) + f[2][k];
        if (qout_cell) qout_cell[gi] = (q[0] + q[1]) + q[2];
        if (qin_cell) qin_cell[gi] = (qi[0] + qi[1]) + qi[2];
    }
}

/* compute_source_terms (source_terms.cuf:4-213) */
void or_source_terms(const or_params *p, int nelt, const double *vx,
                     const double *vy, const double *surf, const double *zm,
                     const double *manning, const int *b1, const double *U0,
                     double dt, double *sb, double *sf, double *afm1)
{
    const double g = p->gp, tls = p->tolisec, tol = p->tol;
    int gi;
#pragma omp parallel for schedule(static)
    for (gi = 0; gi < nelt; gi++) {
        double q1 = U0[3 * gi], q2 = U0[3 * gi + 1], q3 = U0[3 * gi + 2];
        double area = surf[gi];
        double mann = p->is_override_manning ? p->override_manning : manning[gi];
        double s2 = 0.0, s3 = 0.0, vitu = 0.0, vitv = 0.0, normvit = 0.0;
        double *a = afm1 + 9 * gi;
        sf[3 * gi] = 0.0; sf[3 * gi + 1] = 0.0; sf[3 * gi + 2] = 0.0;
        for (int k = 0; k < 3; k++) {
            int kv = b1[3 * gi + k];
            if (kv > 0) {
                double n1, n2, ln, zij, hijet;
                or_face_geom(vx, vy, gi, k, &n1, &n2, &ln);
                zij = dmax(zm[gi], zm[kv - 1]);
                hijet = dmax(0.0, (q1 + zm[gi]) - zij);
                s2 = s2 + (((g / 2.0) * (hijet * hijet - q1 * q1)) * ln) * n1;
                s3 = s3 + (((g / 2.0) * (hijet * hijet - q1 * q1)) * ln) * n2;
            }
        }
        sb[3 * gi] = 0.0; sb[3 * gi + 1] = s2; sb[3 * gi + 2] = s3;
        if (p->friction == 1) {
            if (q1 <= tls) {
                sf[3 * gi + 1] = 0.0; sf[3 * gi + 2] = 0.0;
            } else {
                double p13;
                vitu = q2 / q1;
                vitv = q3 / q1;
                normvit = sqrt(vitu * vitu + vitv * vitv);
                if (normvit < tol) normvit = tol;
                p13 = or_pow(q1, -1.0 / 3.0);
                sf[3 * gi + 1] = -(((((g * mann) * mann) * vitu) * normvit) * area * p13);
                sf[3 * gi + 2] = -(((((g * mann) * mann) * vitv) * normvit) * area * p13);
            }
            for (int k = 0; k < 9; k++) a[k] = (k % 4 == 0) ? 1.0 : 0.0;
            if (p->fricimplic == 1) {
                if (normvit > tol && q1 > tls) {
                    double dtgn2h = ((((dt / 2.0) * g) * mann) * mann) * or_pow(q1, -4.0 / 3.0);
                    double a21 = -((((7.0 / 3.0) * dtgn2h) * vitu) * normvit);
                    double a22 = 1.0 + (dtgn2h * (2.0 * (vitu * vitu) + vitv * vitv)) / normvit;
                    double a23 = ((dtgn2h * vitu) * vitv) / normvit;
                    double a31 = -((((7.0 / 3.0) * dtgn2h) * vitv) * normvit);
                    double a32 = a23;
                    double a33 = 1.0 + (dtgn2h * (vitu * vitu + 2.0 * (vitv * vitv))) / normvit;
                    double det = a22 * a33 - a23 * a32;
                    a[3] = -(a21 * a33 - a23 * a31) / det;
                    a[4] = a33 / det;
                    a[5] = -a23 / det;
                    a[6] = (a21 * a32 - a22 * a31) / det;
                    a[7] = -a32 / det;
                    a[8] = a22 / det;
                }
            } else if (p->fricimplic == 2) {
                if (normvit > tol && q1 > tls) {
                    double beta = -(((((dt * g) * mann) * mann) * normvit) * or_pow(q1, -4.0 / 3.0));
                    double det = (1.0 - beta) * (1.0 - beta);
                    a[4] = (1.0 - beta) / det;
                    a[8] = (1.0 - beta) / det;
                }
            }
        }
    }
}

/* update_sol_with_flux_and_source (finite_volumes_method.cuf:260-305) */
void or_update(const or_params *p, int nelt, const double *surf,
               const double *U0, const double *resc, const double *sb,
               const double *sf, const double *afm1, double dt, double *U)
{
    int gi;
#pragma omp parallel for schedule(static)
    for (gi = 0; gi < nelt; gi++) {
        double res[3];
        if (p->friction == 1) {
            double r[3];
            for (int k = 0; k < 3; k++) r[k] = ((sb[3 * gi + k] + sf[3 * gi + k]) - resc[3 * gi + k]) / surf[gi];
            if (p->fricimplic == 1 || p->fricimplic == 2) {
                const double *a = afm1 + 9 * gi;
                res[0] = (a[0] * r[0] + a[1] * r[1]) + a[2] * r[2];
                res[1] = (a[3] * r[0] + a[4] * r[1]) + a[5] * r[2];
                res[2] = (a[6] * r[0] + a[7] * r[1]) + a[8] * r[2];
            } else {
                res[0] = r[0]; res[1] = r[1]; res[2] = r[2];
            }
        } else {
            for (int k = 0; k < 3; k++) res[k] = (sb[3 * gi + k] - resc[3 * gi + k]) / surf[gi];
        }
        for (int k = 0; k < 3; k++) U[3 * gi + k] = U0[3 * gi + k] + dt * res[k];
        if (U[3 * gi] <= p->tolisec) { U[3 * gi] = 0.0; U[3 * gi + 1] = 0.0; U[3 * gi + 2] = 0.0; }
    }
}

/* cflcond + reduction (cfl_condition.cuf:4-39, :68-104): the minimum over
 * cells of the CFL time step, with the reset value 1e6 and the 1e5 padding
 * of the last 1024-wide reduction block. */
double or_cfl_dt(const or_params *p, int nelt, const double *cotemin_arr, const double *U)
{
    double dt = 1000000.0;
    int gi;
    if (nelt % 1024 != 0) dt = dmin(dt, 100000.0);
#pragma omp parallel for reduction(min : dt) schedule(static)
    for (gi = 0; gi < nelt; gi++) {
        double d = 1000000.0, h = U[3 * gi];
        if (h > p->tolisec) {
            double u = U[3 * gi + 1] / h, v = U[3 * gi + 2] / h;
            double nv = sqrt(u * u + v * v);
            double vmax = dmax(0.01, nv + sqrt(p->gp * fabs(h)));
            d = dmin((p->cfl * cotemin_arr[gi]) / vmax, 1000000.0);
        }
        if (p->is_dt_constant == 1) d = p->constant_dt;
        if (d < dt) dt = d;
    }
    return dt;
}

/* calcul_surf_entree (maillage.cuf:1368-1421): for each inlet id, the sum
 * over the cells with a face on it of that face's length times the cell's
 * depth.  A cell with several faces on the inlet counts its last one (the
 * reference assigns, not adds, per face).  Summed over the cells in cell
 * order: the reference adds block sums with atomics, in no fixed order. */
void or_surf_entree(const or_params *p, int nelt, const double *vx, const double *vy, const int *b1,
                    const int *b2, const double *U, double *surf_entree)
{
    for (int id = 1; id <= OR_MAX_IO; id++) {
        double s = 0.0;
        for (int gi = 0; gi < nelt; gi++) {
            double c = 0.0;
            int hit = 0;
            for (int tj = 0; tj < 3; tj++)
                if (b1[3 * gi + tj] == -1 && b2[3 * gi + tj] == id) {
                    double n1, n2, ln;
                    or_face_geom(vx, vy, gi, tj, &n1, &n2, &ln);
                    c = ln * U[3 * gi];
                    hit = 1;
                }
            if (hit) s = s + c;
        }
        surf_entree[id - 1] = s;
    }
    (void)p;
}

/* solve_finite_volumes time loop (finite_volumes_method.cuf:113-242) */
int or_run(const or_params *p_in, int nelt, const int *conn, const double *vx,
           const double *vy, const double *xc, const double *yc,
           const double *surf, const double *zm, const double *cotemin_arr,
           const double *manning, const int *b1, const int *b2, const int *upw,
           double *U, int nsteps, double *dt_hist, double *tc_out,
           int *steps_done, double *qout_hist, double *qin_hist, int *steady, int nthreads)
{
    size_t n3 = (size_t)nelt * 3;
    double *U1 = (double *)malloc(n3 * sizeof(double));
    double *resc = (double *)malloc(n3 * sizeof(double));
    double *sb = (double *)malloc(n3 * sizeof(double));
    double *sf = (double *)malloc(n3 * sizeof(double));
    double *afm1 = (double *)malloc((size_t)nelt * 9 * sizeof(double));
    double *qc = (double *)malloc((size_t)nelt * sizeof(double));
    double *qic = (double *)malloc((size_t)nelt * sizeof(double));
    double tc = 0.0, dt, dtn;
    int n = 0, status = 0, reg = 0, reached = 0;
    or_params pl = *p_in;  /* the per-step inlet areas live in this copy */
    const or_params *p = &pl;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    for (size_t i = 0; i < (size_t)nelt * 9; i++) afm1[i] = (i % 9) % 4 == 0 ? 1.0 : 0.0;
    /* the initial dt is not range-checked (finite_volumes_method.cuf:113-116) */
    dt = or_cfl_dt(p, nelt, cotemin_arr, U);
    while (status == 0 && !reached && n < nsteps && ((tc - p->ts) + 2.0 * p->tol) < p->tol) {
        double qo = 0.0, qi = 0.0;
        if (p->inlet == 3)  /* finite_volumes_method.cuf:133-139 */
            or_surf_entree(p, nelt, vx, vy, b1, b2, U, pl.surf_entree);
        or_compute_flux(p, nelt, conn, vx, vy, xc, yc, zm, b1, b2, upw, U, dt, tc, resc, qc, qic);
        or_source_terms(p, nelt, vx, vy, surf, zm, manning, b1, U, dt, sb, sf, afm1);
        or_update(p, nelt, surf, U, resc, sb, sf, afm1, dt, U1);
        for (int i = 0; i < nelt; i++) {
            qo = qo + qc[i];
            qi = qi + qic[i];
        }
        if (qout_hist) qout_hist[n] = qo;
        if (qin_hist) qin_hist[n] = qi;
        dtn = or_cfl_dt(p, nelt, cotemin_arr, U1);
        memcpy(U, U1, n3 * sizeof(double));
        dt_hist[n] = dt;
        tc = tc + dt;
        n++;
        dt = dtn;
        /* check_steady_state (finite_volumes_method.cuf:370-393) on this
           step's discharges: more than 1000 consecutive steps with
           |Qin - Qout| / Qin < tol_reg_perm end the run after the step */
        if (qi > LIT_1EM8 && fabs(qi - qo) / qi < p->tol_reg_perm) {
            if (++reg > 1000) reached = 1;
        } else {
            reg = 0;
        }
        /* dt range check of cfl_condition.cuf:129-132 applied to the new dt;
           NaN states are an error too (the reference does not catch them) */
        if (!(dt <= 9999.0 && dt >= LIT_1EM12)) status = -1;
        for (size_t i = 0; i < n3 && status == 0; i++)
            if (U[i] != U[i]) status = -1;
    }
    dt_hist[n] = dt;
    if (tc_out) *tc_out = tc;
    if (steps_done) *steps_done = n;
    if (steady) *steady = reached;
    free(U1); free(resc); free(sb); free(sf); free(afm1); free(qc); free(qic);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Direct access to the face solvers and the limiter for the pin tests.     */

void or_riemann_face(int iflux, int ilimiteur, double gp, double n1, double n2,
                     double hl, double ul, double vl, double hr, double ur, double vr,
                     double hl_upw, double vl_upw, double hr_upw, double vr_upw,
                     double d_mrl, double d_mkvol, double d_mkvoisin, double dt,
                     double flux[3])
{
    if (iflux == 1) {
        or_hllc(gp, n1, n2, hl, ul, vl, hr, ur, vr, flux);
    } else if (iflux == 2) {
        or_riadh(gp, n1, n2, hl, ul, vl, hr, ur, vr, flux);
    } else {
        or_params p;
        memset(&p, 0, sizeof(p));
        p.gp = gp;
        p.ilimiteur = ilimiteur;
        or_waf(&p, n1, n2, hl, ul, vl, hr, ur, vr, hl_upw, vl_upw, hr_upw, vr_upw,
               d_mrl, d_mkvol, d_mkvoisin, dt, flux);
    }
}

double or_limiter_value(int ilimiteur, double r, double c)
{
    return or_limiter(ilimiteur, r, c);
}

double or_slope_limiter_value(int ilimiteur, double r1, double r2)
{
    return or_slope_limiter(ilimiteur, r1, r2);
}

void or_muscl_face(int ilimiteur, double gp, double tolisec, double n1, double n2,
                   double hl, double ul, double vl, double hr, double ur, double vr,
                   double hl_upw, double ul_upw, double vl_upw,
                   double hr_upw, double ur_upw, double vr_upw,
                   double d_mrl, double d_mkvol, double d_mkvoisin, double dt, double flux[3])
{
    or_params p;
    memset(&p, 0, sizeof(p));
    p.gp = gp;
    p.tolisec = tolisec;
    p.ilimiteur = ilimiteur;
    or_muscl(&p, n1, n2, hl, ul, vl, hr, ur, vr, hl_upw, ul_upw, vl_upw, hr_upw, ur_upw, vr_upw,
             d_mrl, d_mkvol, d_mkvoisin, dt, flux);
}
