// CPU build of the step's cell body (cuteflow_amd/csrc/cf_cell.h) over a whole
// mesh: forward, tangent-linear and reverse sweeps, for checks of the
// derivative code without a GPU.  Test infrastructure only: the library never
// runs this code.  Cells keep the caller's order (no Hilbert reordering).
#include <cmath>
#include <cstring>
#include <vector>

#include "../../cuteflow_amd/csrc/cf_cell.h"
#include "../../cuteflow_amd/csrc/cf_prep.h"

using namespace cf;

namespace {

struct HLd {
    const double *st;
    const I4 *nb;
    const D4 *df;
    const double *g;
    const double *man;
    int N;
    St<double> state(int c) const { return {st[4 * c], st[4 * c + 1], st[4 * c + 2], st[4 * c + 3]}; }
    I4 nbr(int c) const { return nb[c]; }
    D4 dfc(int c) const { return df[c]; }
    double geo(int comp, int c) const { return g[(size_t)comp * N + c]; }
    double manning(int c) const { return man[c]; }
};

struct HLdT {
    const double *st, *dst;
    const I4 *nb;
    const D4 *df;
    const double *g;
    const double *man, *dman;
    int N;
    St<Dual> state(int c) const {
        return {Dual(st[4 * c], dst[4 * c]), Dual(st[4 * c + 1], dst[4 * c + 1]), Dual(st[4 * c + 2], dst[4 * c + 2]),
                st[4 * c + 3]};
    }
    I4 nbr(int c) const { return nb[c]; }
    D4 dfc(int c) const { return df[c]; }
    double geo(int comp, int c) const { return g[(size_t)comp * N + c]; }
    Dual manning(int c) const { return Dual(man[c], dman[c]); }
};

struct HS {
    int N = 0, iflux = 1, fric = 0;
    KParams kp{};
    std::vector<double> hydro;
    std::vector<I4> nb;
    std::vector<D4> df;
    std::vector<double> geo, man, zm;
    int outlet = 1;
    // recorded run
    int nsteps = 0;
    std::vector<double> dt, tc;
    std::vector<std::vector<double>> U, X;  // U[k]: 4N, X[k]: 3N
};

template <int IF, int FR>
void step_t(const HS &s, const double *Uin, double *Uout, double *X, double dt, double tc) {
    const HLd ld{Uin, s.nb.data(), s.df.data(), s.geo.data(), s.man.data(), s.N};
    for (int i = 0; i < s.N; i++) {
        double o[3], x[3];
        cell_update<double, IF, FR>(ld, s.kp, s.hydro.data(), i, dt, tc, o, x);
        Uout[4 * i] = o[0];
        Uout[4 * i + 1] = o[1];
        Uout[4 * i + 2] = o[2];
        Uout[4 * i + 3] = Uin[4 * i + 3];
        if (X) {
            X[i] = x[0];
            X[(size_t)s.N + i] = x[1];
            X[2 * (size_t)s.N + i] = x[2];
        }
    }
}

template <int IF, int FR>
void tlm_t(const HS &s, const double *Uin, const double *dUin, const double *dman, double *Uout, double *dUout,
           double dt, double tc) {
    const HLdT ld{Uin, dUin, s.nb.data(), s.df.data(), s.geo.data(), s.man.data(), dman, s.N};
    for (int i = 0; i < s.N; i++) {
        Dual o[3];
        cell_update<Dual, IF, FR>(ld, s.kp, s.hydro.data(), i, dt, tc, o, (Dual *)nullptr);
        for (int k = 0; k < 3; k++) {
            Uout[4 * i + k] = o[k].v;
            dUout[4 * i + k] = o[k].d;
        }
        Uout[4 * i + 3] = Uin[4 * i + 3];
        dUout[4 * i + 3] = 0.0;
    }
}

// Second-ring target of face j, slot s of cell i (the cell the WAF upwind
// state of that slot is read from), or -1 when that slot is not read.
int r2_target(const HS &s, int i, int j, int slot) {
    const I4 v = s.nb[i];
    const int c[3] = {v.x, v.y, v.z};
    int kv = -1;
    if (c[j] >= 0) kv = c[j];
    else if (((-c[j]) & 3) == 2 && s.outlet == 2) kv = i;
    if (kv < 0) return -1;
    const bool own1 = c[(j + 1) % 3] < 0, own2 = c[(j + 2) % 3] < 0;
    const int tjv = tjv_of(v, j);
    const I4 vk = (kv == i) ? v : s.nb[kv];
    const int e1 = nb_at(vk, (tjv + 1) % 3), e2 = nb_at(vk, (tjv + 2) % 3);
    const int kw1 = e1 >= 0 ? e1 : kv, kw2 = e2 >= 0 ? e2 : kv;
    if ((own1 && own2) || (kw1 == kv && kw2 == kv)) return -1;
    const int kw = slot == 0 ? kw1 : kw2;
    return kw != kv ? kw : -1;
}

// Returns the number of nonzero cotangents addressed to slots that the
// forward does not read (must be 0).
template <int IF, int FR>
long rev_t(const HS &s, const double *Uin, const double *X, const double *lam, double *lam_prev, double *gn,
           double dt, double tc) {
    const HLd ld{Uin, s.nb.data(), s.df.data(), s.geo.data(), s.man.data(), s.N};
    long bad = 0;
    for (int i = 0; i < s.N; i++) {
        const double x[3] = {X[i], X[(size_t)s.N + i], X[2 * (size_t)s.N + i]};
        const double l[3] = {lam[3 * i], lam[3 * i + 1], lam[3 * i + 2]};
        double own[3], r1[3][3], r2[3][2][3], g;
        cell_reverse<IF, FR>(ld, s.kp, s.hydro.data(), i, dt, tc, x, l, own, r1, r2, g);
        for (int k = 0; k < 3; k++) lam_prev[3 * i + k] += own[k];
        const I4 v = s.nb[i];
        for (int j = 0; j < 3; j++) {
            const int t = nb_at(v, j);
            for (int k = 0; k < 3; k++) {
                if (t >= 0) lam_prev[3 * t + k] += r1[j][k];
                else if (r1[j][k] != 0.0) bad++;
            }
            for (int sl = 0; sl < 2; sl++) {
                const int tg = r2_target(s, i, j, sl);
                for (int k = 0; k < 3; k++) {
                    if (tg >= 0) lam_prev[3 * tg + k] += r2[j][sl][k];
                    else if (r2[j][sl][k] != 0.0) bad++;
                }
            }
        }
        gn[i] += g;
    }
    return bad;
}

typedef void (*step_fn)(const HS &, const double *, double *, double *, double, double);
typedef void (*tlm_fn)(const HS &, const double *, const double *, const double *, double *, double *, double, double);
typedef long (*rev_fn)(const HS &, const double *, const double *, const double *, double *, double *, double, double);

step_fn pick_step(int iflux, int fric) {
    static const step_fn t[4][4] = {{step_t<1, 0>, step_t<1, 1>, step_t<1, 2>, step_t<1, 3>},
                                  {step_t<2, 0>, step_t<2, 1>, step_t<2, 2>, step_t<2, 3>},
                                  {step_t<3, 0>, step_t<3, 1>, step_t<3, 2>, step_t<3, 3>},
                                  {step_t<4, 0>, step_t<4, 1>, step_t<4, 2>, step_t<4, 3>}};
    return t[iflux - 1][fric];
}
tlm_fn pick_tlm(int iflux, int fric) {
    static const tlm_fn t[4][4] = {{tlm_t<1, 0>, tlm_t<1, 1>, tlm_t<1, 2>, tlm_t<1, 3>},
                                  {tlm_t<2, 0>, tlm_t<2, 1>, tlm_t<2, 2>, tlm_t<2, 3>},
                                  {tlm_t<3, 0>, tlm_t<3, 1>, tlm_t<3, 2>, tlm_t<3, 3>},
                                  {tlm_t<4, 0>, tlm_t<4, 1>, tlm_t<4, 2>, tlm_t<4, 3>}};
    return t[iflux - 1][fric];
}
rev_fn pick_rev(int iflux, int fric) {
    static const rev_fn t[4][4] = {{rev_t<1, 0>, rev_t<1, 1>, rev_t<1, 2>, rev_t<1, 3>},
                                  {rev_t<2, 0>, rev_t<2, 1>, rev_t<2, 2>, rev_t<2, 3>},
                                  {rev_t<3, 0>, rev_t<3, 1>, rev_t<3, 2>, rev_t<3, 3>},
                                  {rev_t<4, 0>, rev_t<4, 1>, rev_t<4, 2>, rev_t<4, 3>}};
    return t[iflux - 1][fric];
}

// cflcond + reduction (cfl_condition.cuf:4-104): dt of a state
double cfl_dt(const HS &s, const double *U) {
    double m = s.kp.pad_dt;
    for (int i = 0; i < s.N; i++) {
        double dti = 1000000.0;
        const double h = U[4 * i];
        if (h > s.kp.tolisec) {
            const double uu = U[4 * i + 1] / h, vv = U[4 * i + 2] / h;
            const double nv = sqrt(uu * uu + vv * vv);
            const double vmax = dmax(0.01, nv + sqrt(s.kp.gp * fabs(h)));
            dti = dmin((s.kp.cfl * s.geo[(size_t)G_COT * s.N + i]) / vmax, 1000000.0);
        }
        if (s.kp.is_dt_constant == 1) dti = s.kp.constant_dt;
        m = dmin(m, dti);
    }
    return m;
}

}  // namespace

extern "C" {

void *hs_create(const cf_mesh *m, const cf_params *p) {
    cfprep::Prep prep;
    if (cfprep::prepare(m, p, 256, true, prep) != 0) return nullptr;
    HS *s = new HS();
    const int N = m->nelt;
    s->N = N;
    s->iflux = p->iflux;
    s->fric = p->friction == 0 ? 0 : 1 + p->fricimplic;
    s->outlet = p->outlet;
    KParams &k = s->kp;
    k.gp = p->gp; k.tolisec = p->tolisec; k.tol = p->tol; k.cfl = p->cfl; k.ts = p->ts;
    k.constant_dt = p->constant_dt; k.override_manning = p->override_manning;
    k.pad_dt = (N % 1024 != 0) ? 100000.0 : 1000000.0;
    k.ilimiteur = p->ilimiteur; k.inlet = p->inlet; k.outlet = p->outlet;
    k.is_dt_constant = p->is_dt_constant; k.is_override_manning = p->is_override_manning;
    k.hydro_off = 0;
    for (int i = 0; i < CF_MAX_IO; i++) {
        k.debitglob[i] = p->debitglob[i];
        k.h_sortie[i] = p->h_sortie[i];
        k.long_entree[i] = p->long_entree[i];
        k.hydro_n[i] = p->hydro_n[i];
    }
    s->hydro.resize(2 * CF_MAX_IO * CF_MAX_HYDRO);
    memcpy(s->hydro.data(), p->hydro_t, sizeof(double) * CF_MAX_IO * CF_MAX_HYDRO);
    memcpy(s->hydro.data() + CF_MAX_IO * CF_MAX_HYDRO, p->hydro_q, sizeof(double) * CF_MAX_IO * CF_MAX_HYDRO);
    s->nb.resize(N);
    for (int i = 0; i < N; i++) s->nb[i] = {prep.nbr[i].c[0], prep.nbr[i].c[1], prep.nbr[i].c[2], prep.nbr[i].w};
    s->geo.assign((size_t)G_COUNT * N, 0.0);
    s->zm.assign(N, 0.0);
    for (int i = 0; i < N; i++) {
        double vx[3], vy[3], vz[3];
        for (int a = 0; a < 3; a++) {
            const int n = prep.conn[3 * i + a] - 1;
            vx[a] = m->coords[3 * n];
            vy[a] = m->coords[3 * n + 1];
            vz[a] = m->coords[3 * n + 2];
        }
        double g[G_COUNT];
        cell_geometry(vx, vy, vz, g, s->zm[i]);
        for (int c = 0; c < G_COUNT; c++) s->geo[(size_t)c * N + i] = g[c];
    }
    s->df.resize(N);
    for (int i = 0; i < N; i++) {
        double d[3];
        const double xi = s->geo[(size_t)G_XC * N + i], yi = s->geo[(size_t)G_YC * N + i];
        for (int j = 0; j < 3; j++) {
            const int kk = nb_at(s->nb[i], j) >= 0 ? nb_at(s->nb[i], j) : i;
            const double xk = s->geo[(size_t)G_XC * N + kk], yk = s->geo[(size_t)G_YC * N + kk];
            d[j] = sqrt((xi - xk) * (xi - xk) + (yi - yk) * (yi - yk));
        }
        s->df[i] = {d[0], d[1], d[2], 0.0};
    }
    s->man.assign(N, 0.0);
    if (m->manning)
        for (int i = 0; i < N; i++) s->man[i] = m->manning[i];
    return s;
}

void hs_destroy(void *h) { delete (HS *)h; }

void hs_set_ad_flags(void *h, int flags) { ((HS *)h)->kp.ad_flags = flags; }

void hs_set_manning(void *h, const double *n) {
    HS *s = (HS *)h;
    for (int i = 0; i < s->N; i++) s->man[i] = n[i];
}

// Forward run of nsteps steps from U0 (3N), recording every state: dt_hist
// (nsteps+1) and the final state (3N).  Returns the steps run (a step whose
// new dt is out of range or NaN ends the run).
int hs_forward(void *h, int nsteps, const double *U0, double *dt_hist, double *Ufinal) {
    HS *s = (HS *)h;
    const int N = s->N;
    s->U.assign(nsteps + 1, std::vector<double>(4 * (size_t)N));
    s->X.assign(nsteps, std::vector<double>(3 * (size_t)N));
    s->dt.assign(nsteps + 1, 0.0);
    s->tc.assign(nsteps + 1, 0.0);
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < 3; k++) s->U[0][4 * i + k] = U0[3 * i + k];
        s->U[0][4 * i + 3] = s->zm[i];
    }
    s->dt[0] = cfl_dt(*s, s->U[0].data());
    const step_fn f = pick_step(s->iflux, s->fric);
    int k = 0;
    for (; k < nsteps; k++) {
        f(*s, s->U[k].data(), s->U[k + 1].data(), s->X[k].data(), s->dt[k], s->tc[k]);
        s->tc[k + 1] = s->tc[k] + s->dt[k];
        const double d = cfl_dt(*s, s->U[k + 1].data());
        s->dt[k + 1] = d;
        if (!(d <= 9999.0 && d >= kDtLow)) {
            k++;
            break;
        }
    }
    s->nsteps = k;
    for (int j = 0; j <= k; j++) dt_hist[j] = s->dt[j];
    for (int i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) Ufinal[3 * i + c] = s->U[k][4 * i + c];
    return k;
}

// Tangent-linear run along (dn, dU0) over the recorded run.
void hs_tlm(void *h, const double *dn, const double *dU0, double *dUfinal, double *Ufinal) {
    HS *s = (HS *)h;
    const int N = s->N;
    std::vector<double> u(s->U[0]), du(4 * (size_t)N, 0.0), u2(4 * (size_t)N), du2(4 * (size_t)N);
    for (int i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) du[4 * i + c] = dU0 ? dU0[3 * i + c] : 0.0;
    std::vector<double> dm(N, 0.0);
    if (dn)
        for (int i = 0; i < N; i++) dm[i] = dn[i];
    const tlm_fn f = pick_tlm(s->iflux, s->fric);
    for (int k = 0; k < s->nsteps; k++) {
        f(*s, u.data(), du.data(), dm.data(), u2.data(), du2.data(), s->dt[k], s->tc[k]);
        u.swap(u2);
        du.swap(du2);
    }
    for (int i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) {
            dUfinal[3 * i + c] = du[4 * i + c];
            Ufinal[3 * i + c] = u[4 * i + c];
        }
}

// Reverse sweep over the recorded run from the final-state cotangent:
// gradients with respect to Manning n (N) and the initial state (3N).
long hs_adjoint(void *h, const double *lam_final, double *g_n, double *g_U0) {
    HS *s = (HS *)h;
    const int N = s->N;
    std::vector<double> lam(lam_final, lam_final + 3 * (size_t)N), prev(3 * (size_t)N);
    std::vector<double> gn(N, 0.0);
    const rev_fn f = pick_rev(s->iflux, s->fric);
    long bad = 0;
    for (int k = s->nsteps - 1; k >= 0; k--) {
        std::fill(prev.begin(), prev.end(), 0.0);
        bad += f(*s, s->U[k].data(), s->X[k].data(), lam.data(), prev.data(), gn.data(), s->dt[k], s->tc[k]);
        lam.swap(prev);
    }
    for (int i = 0; i < N; i++) g_n[i] = gn[i];
    for (size_t i = 0; i < 3 * (size_t)N; i++) g_U0[i] = lam[i];
    return bad;
}

// Face-level check: value of <bflux, J dir> by the dual-number forward and
// <J^T bflux, dir> by the reverse, for the first-order (iflux 1 and 2,
// inputs 0..5) or WAF (iflux 3, inputs 0..9) face flux in the global frame.
void hs_face_dot(int iflux, int ilim, double gp, double n1, double n2, const double *in, const double *dir,
                 const double *bflux, double d_mrl, double d_mkvol, double d_mkvoisin, double dt, double *out) {
    Dual v[10];
    for (int k = 0; k < 10; k++) v[k] = Dual(in[k], dir[k]);
    Dual fl[3];
    double g[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (iflux == 1) {
        hllc_flux(gp, n1, n2, v[0], v[1], v[2], v[3], v[4], v[5], fl);
        hllc_flux_rev(gp, n1, n2, in[0], in[1], in[2], in[3], in[4], in[5], bflux, g);
    } else if (iflux == 2) {
        riadh_flux(gp, n1, n2, v[0], v[1], v[2], v[3], v[4], v[5], fl);
        riadh_flux_rev(gp, n1, n2, in[0], in[1], in[2], in[3], in[4], in[5], bflux, g);
    } else {
        waf_flux(gp, ilim, n1, n2, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], d_mrl, d_mkvol,
                 d_mkvoisin, dt, fl);
        waf_flux_rev(gp, ilim, n1, n2, in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9], d_mrl,
                     d_mkvol, d_mkvoisin, dt, bflux, g);
    }
    double fwd = 0.0, rev = 0.0;
    for (int k = 0; k < 3; k++) fwd += bflux[k] * fl[k].d;
    for (int k = 0; k < 10; k++) rev += g[k] * dir[k];
    out[0] = fwd;
    out[1] = rev;
    out[2] = fl[0].v;
    out[3] = fl[1].v;
    out[4] = fl[2].v;
}

// The same check for the MUSCL face flux (iflux 4): inputs (hl, ul, vl, hr,
// ur, vr, hl_upw, ul_upw, vl_upw, hr_upw, ur_upw, vr_upw).
void hs_muscl_face_dot(int ilim, double gp, double tolisec, double n1, double n2, const double *in, const double *dir,
                       const double *bflux, double d_mrl, double d_mkvol, double d_mkvoisin, double dt, double *out) {
    Dual v[12];
    for (int k = 0; k < 12; k++) v[k] = Dual(in[k], dir[k]);
    Dual fl[3];
    double g[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    muscl_flux(gp, ilim, tolisec, n1, n2, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10], v[11],
               d_mrl, d_mkvol, d_mkvoisin, dt, fl);
    muscl_flux_rev(gp, ilim, tolisec, n1, n2, in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9],
                   in[10], in[11], d_mrl, d_mkvol, d_mkvoisin, dt, bflux, g);
    double fwd = 0.0, rev = 0.0;
    for (int k = 0; k < 3; k++) fwd += bflux[k] * fl[k].d;
    for (int k = 0; k < 12; k++) rev += g[k] * dir[k];
    out[0] = fwd;
    out[1] = rev;
    out[2] = fl[0].v;
    out[3] = fl[1].v;
    out[4] = fl[2].v;
}

}  // extern "C"
