/*
 * sld_model_oracle.c - the parametric potential model restated on the CPU.
 * TEST INFRASTRUCTURE ONLY: the checker of libspilady_hip.so's model and
 * adjoint entry points, never the product.
 *
 * Physics of one frozen configuration: the density of calculate_rho_CPU.cpp
 * :33-66, the energy and force of calculate_force_energy_CPU.cpp:33-216 and
 * the effective field of calculate_spin_CPU.cpp:111-140,201-274, with the
 * DD05 (DD05.cpp:6-164) and Fe exchange/Landau (JijFe.cpp:5-43) forms taking
 * their coefficients from theta and tabulated as potential_CPU.cpp:61-170
 * (10^6 samples, 6-point central derivative, linear interpolation).
 *
 * Gradient: FORWARD mode, one parameter at a time. For each theta_k the
 * derivative of every intermediate (rho_i, F'(rho_i), pair terms) is carried
 * through the same formulas as the values, giving dE/dtheta_k, dF_i/dtheta_k
 * and dH_i/dtheta_k explicitly before they are contracted with the upstream
 * gradients. This is a different route from a reverse-mode adjoint; both are
 * exact for the tabulated model (the derivative of a linearly interpolated
 * table in its argument is the slope of the interpolating segment).
 *
 * Pinned by: the tables at the reference's theta against the oracle's own
 * tables (themselves pinned against the reference's golden tables); the
 * derivative generators against central differences in theta; the bridge
 * coefficients against the six end conditions that define them; forces,
 * energies and fields at the reference's theta against the oracle's force
 * and spin paths.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sld_oracle.h"

#define NIN 1000000

/* knots of the density and pair splines (DD05.cpp:26-37,70-85) */
static const double KF[6] = {3.0, 2.8666666666666666, 2.7333333333333333, 2.6, 2.4, 2.3};
static const double KV[8] = {4.1, 3.8, 3.5, 3.2, 2.9, 2.6, 2.4, 2.3};

/* DD05.cpp:6-164, JijFe.cpp:5-43 */
void orcm_default_params(double* th) {
    static const double a[6] = {0.9332056681088162, -1.162558782567700, -0.3502026949249225,
                                0.4287820835430028, 4.907925057809273, -5.307049068415304};
    static const double v[8] = {-0.1960674387419232, 0.3687525935422963, -1.505333614924853,
                                4.948907078156191,   -4.894613262753399, 3.468897724782442,
                                -1.792218099820337,  80.22069592246987};
    for (int k = 0; k < 6; ++k) th[k] = a[k];
    for (int k = 0; k < 8; ++k) th[6 + k] = v[k];
    th[14] = 4.100199340884814;
    th[15] = 1.565647547483517;
    th[16] = 0.870;
    th[17] = 3.75;
    th[18] = -0.744824;
    th[19] = 0.345295;
    th[20] = -0.00790205;
    th[21] = 0e0;
}

/* ---- generators at theta ---- */
static double embed(const double* th, double rho) {
    double aa = th[14], bb = th[15];
    if (rho >= 1e0) return -aa * sqrt(rho);
    return -aa * sqrt(rho) - bb / log(2e0) * (1e0 - sqrt(rho)) * log(2e0 - rho);
}
static double dens(const double* th, double r) {
    double f = 0e0;
    for (int k = 0; k < 6; ++k)
        if (r < KF[k]) f += th[k] * pow(KF[k] - r, 3);
    return f;
}
static double exch(const double* th, double r) {
    double A = th[16], rc = th[17];
    return r < rc ? A * pow(1.0 - r / rc, 3) : 0e0;
}

static double pair_basis(int k, double r);

/* V(r; theta): the reference's closed-form bridge (orc_pair_v) is linear in
 * its end values, so on (1.8, 2] it equals the bridge of the ZBL end alone
 * (orc_pair_v with V = 0) plus sum_k V_k times the bridge of knot k alone;
 * the split keeps the closed form's rounding noise out of the V dependence */
static double pair(const double* th, double r) {
    static const double zero[8] = {0e0, 0e0, 0e0, 0e0, 0e0, 0e0, 0e0, 0e0};
    if (r <= 1.8 || r > 2.0) return orc_pair_v(r, th + 6);
    double v = orc_pair_v(r, zero);
    for (int k = 0; k < 8; ++k) v += th[6 + k] * pair_basis(k, r);
    return v;
}

double orcm_gen(const double* th, int kind, double x) {
    switch (kind) {
        case ORCM_EMBED: return embed(th, x);
        case ORCM_DENS: return dens(th, x);
        case ORCM_PAIR: return pair(th, x);
        case ORCM_EXCH: return exch(th, x);
    }
    return NAN;
}

/* ---- derivatives of the generators in theta ---- */

/* bridge of V_k alone: the ZBL end does not depend on V, so its end values
 * are zero and the spline end values are those of the single knot term */
static double vb_coef[8][6];
static int vb_ready = 0;
static void pair_basis_init(void) {
    if (vb_ready) return;
    for (int k = 0; k < 8; ++k) {
        double e[6] = {0e0, 0e0, 0e0, 0e0, 0e0, 0e0};
        double x2 = 2.0;
        if (x2 < KV[k]) {
            e[3] = pow(KV[k] - x2, 3);
            e[4] = -3e0 * pow(KV[k] - x2, 2);
            e[5] = 6e0 * (KV[k] - x2);
        }
        orc_bridge_coeffs(e, vb_coef[k]);
    }
    vb_ready = 1;
}
static double pair_basis(int k, double r) {
    pair_basis_init();
    if (r <= 1.8) return 0e0;
    if (r <= 2.0) {
        const double* a = vb_coef[k];
        return a[5] * pow(r, 5) + a[4] * pow(r, 4) + a[3] * pow(r, 3) + a[2] * pow(r, 2) + a[1] * r + a[0];
    }
    return r < KV[k] ? pow(KV[k] - r, 3) : 0e0;
}

double orcm_dgen(const double* th, int kind, int k, double x) {
    pair_basis_init();
    switch (kind) {
        case ORCM_EMBED:
            if (k == 14) return -sqrt(x);
            if (k == 15) return x >= 1e0 ? 0e0 : -(1e0 - sqrt(x)) * log(2e0 - x) / log(2e0);
            return 0e0;
        case ORCM_DENS:
            if (k >= 0 && k < 6) return x < KF[k] ? pow(KF[k] - x, 3) : 0e0;
            return 0e0;
        case ORCM_PAIR:
            if (k >= 6 && k < 14) return pair_basis(k - 6, x);
            return 0e0;
        case ORCM_EXCH: {
            double A = th[16], rc = th[17];
            if (x >= rc) return 0e0;
            if (k == 16) return pow(1.0 - x / rc, 3);
            if (k == 17) return A * 3.0 * pow(1.0 - x / rc, 2) * x / (rc * rc);
            return 0e0;
        }
    }
    return NAN;
}

/* potential_CPU.cpp:159-170: 6-point central difference, dx = x/1000 */
static double d6(const double* th, int kind, int k, double x) {
    double dx = x / 1000e0;
    double v[6];
    const double off[6] = {1, -1, 2, -2, 3, -3};
    for (int q = 0; q < 6; ++q)
        v[q] = k < 0 ? orcm_gen(th, kind, x + off[q] * dx) : orcm_dgen(th, kind, k, x + off[q] * dx);
    return (45e0 * (v[0] - v[1]) - 9e0 * (v[2] - v[3]) + (v[4] - v[5])) / (60e0 * dx);
}

/* ---- tables ---- */
static int kind_of(int k) {
    if (k < 6) return ORCM_DENS;
    if (k < 14) return ORCM_PAIR;
    if (k < 16) return ORCM_EMBED;
    if (k < 18) return ORCM_EXCH;
    return -1;
}

struct orcm_tab {
    double th[ORCM_NPARAM];
    double rmax, rhomax, in_r, in_rho;
    double* val[4][2];               /* [kind][value, derivative] */
    double* der[ORCM_NPARAM][2];     /* d table / d theta_k, NULL for the Landau entries */
};

orcm_tab* orcm_build(const double* th, double rcut_max) {
    pair_basis_init();
    orcm_tab* t = (orcm_tab*)calloc(1, sizeof(orcm_tab));
    memcpy(t->th, th, sizeof(t->th));
    t->rmax = rcut_max;
    t->rhomax = 2e0 * dens(th, 1e0);
    if (t->rhomax < 1000e0) t->rhomax = 1000e0;
    t->in_r = (double)NIN / t->rmax;
    t->in_rho = (double)NIN / t->rhomax;
    for (int q = 0; q < 4; ++q)
        for (int d = 0; d < 2; ++d) t->val[q][d] = (double*)malloc(sizeof(double) * NIN);
    for (int k = 0; k < 18; ++k)
        for (int d = 0; d < 2; ++d) t->der[k][d] = (double*)malloc(sizeof(double) * NIN);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < NIN; ++n) {
        double r = t->rmax / (double)NIN * (double)(n + 1);
        double rho = t->rhomax / (double)NIN * (double)(n + 1);
        for (int q = 0; q < 4; ++q) {
            double x = q == ORCM_EMBED ? rho : r;
            t->val[q][0][n] = orcm_gen(th, q, x);
            t->val[q][1][n] = d6(th, q, -1, x);
        }
        for (int k = 0; k < 18; ++k) {
            int q = kind_of(k);
            double x = q == ORCM_EMBED ? rho : r;
            t->der[k][0][n] = orcm_dgen(th, q, k, x);
            t->der[k][1][n] = d6(th, q, k, x);
        }
    }
    return t;
}

void orcm_free(orcm_tab* t) {
    if (!t) return;
    for (int q = 0; q < 4; ++q)
        for (int d = 0; d < 2; ++d) free(t->val[q][d]);
    for (int k = 0; k < 18; ++k)
        for (int d = 0; d < 2; ++d) free(t->der[k][d]);
    free(t);
}

/* kind >= 0: value table of that kind; kind <= -1: derivative table of
 * parameter -1 - kind */
double orcm_table_raw(const orcm_tab* t, int kind, int deriv, int n) {
    if (n < 0 || n >= NIN || deriv < 0 || deriv > 1) return NAN;
    if (kind >= 0 && kind < 4) return t->val[kind][deriv][n];
    int k = -1 - kind;
    if (k >= 0 && k < 18) return t->der[k][deriv][n];
    return NAN;
}

/* potential_CPU.cpp:172-179, index kept in [1, N-1]; *slope = d value / dx */
static inline double lerp(const double* tb, double x, double inv, double* slope) {
    double fn = x * inv;
    int n = (int)fn;
    if (n < 1) n = 1;
    if (n > NIN - 1) n = NIN - 1;
    double ratio = (double)(n + 1) - fn;
    if (slope) *slope = (tb[n] - tb[n - 1]) * inv;
    return tb[n - 1] * ratio + tb[n] * (1e0 - ratio);
}

/* ---- one configuration ---- */
typedef struct {
    int i, j;
    double d[3], r;
} pair_t;

static double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

int orcm_eval_grad(const orcm_tab* t, const double L[3], int n, const double* r, const double* s,
                   double rcut_max, double rcut_pot, double rcut_mag, double gE, const double* gF,
                   const double* gH, double* E, double* F, double* H, double* grad) {
    for (int d = 0; d < 3; ++d)
        if (L[d] < 2.0 * rcut_max) return -1;
    const double* th = t->th;
    const double rmax2 = rcut_max * rcut_max;
    /* every pair once, i < j, with the minimum image of r_i - r_j */
    size_t cap = (size_t)n * 32, np = 0;
    pair_t* P = (pair_t*)malloc(sizeof(pair_t) * cap);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d[3];
            int far = 0;
            for (int q = 0; q < 3 && !far; ++q) {
                double x = r[3 * i + q] - r[3 * j + q];
                if (x > 0.5 * L[q]) x -= L[q];
                else if (x < -0.5 * L[q]) x += L[q];
                d[q] = x;
                if (fabs(x) >= rcut_max) far = 1;
            }
            if (far) continue;
            double rr = dot(d, d);
            if (rr >= rmax2) continue;
            if (np == cap) {
                cap *= 2;
                P = (pair_t*)realloc(P, sizeof(pair_t) * cap);
            }
            P[np].i = i;
            P[np].j = j;
            memcpy(P[np].d, d, sizeof(d));
            P[np].r = sqrt(rr);
            ++np;
        }
    double* rho = (double*)calloc(n, sizeof(double));
    double* fp = (double*)calloc(n, sizeof(double));
    double* sF = (double*)calloc(n, sizeof(double));
    double* sFp = (double*)calloc(n, sizeof(double));
    double* s0 = (double*)calloc(n, sizeof(double));
    double* drho = (double*)calloc(n, sizeof(double));
    double* dfp = (double*)calloc(n, sizeof(double));
    double* dF = (double*)calloc((size_t)3 * n, sizeof(double));
    double* dH = (double*)calloc((size_t)3 * n, sizeof(double));
    const double *tF = t->val[ORCM_EMBED][0], *tFp = t->val[ORCM_EMBED][1];
    const double *tf = t->val[ORCM_DENS][0], *tfp = t->val[ORCM_DENS][1];
    const double *tV = t->val[ORCM_PAIR][0], *tVp = t->val[ORCM_PAIR][1];
    const double *tJ = t->val[ORCM_EXCH][0], *tJp = t->val[ORCM_EXCH][1];

    /* values */
    for (size_t p = 0; p < np; ++p)
        if (P[p].r < rcut_pot) {
            double f = lerp(tf, P[p].r, t->in_r, NULL);
            rho[P[p].i] += f;
            rho[P[p].j] += f;
        }
    double e = 0e0;
    for (int i = 0; i < n; ++i) {
        fp[i] = lerp(tFp, rho[i], t->in_rho, &sFp[i]);
        e += lerp(tF, rho[i], t->in_rho, &sF[i]);
        s0[i] = sqrt(dot(s + 3 * i, s + 3 * i));
    }
    memset(F, 0, sizeof(double) * 3 * n);
    memset(H, 0, sizeof(double) * 3 * n);
    double* sumJs = (double*)calloc(n, sizeof(double));
    for (size_t p = 0; p < np; ++p) {
        const pair_t* q = &P[p];
        const int i = q->i, j = q->j;
        double dudr = 0e0;
        if (q->r < rcut_pot) {
            dudr += (fp[i] + fp[j]) * lerp(tfp, q->r, t->in_r, NULL) + lerp(tVp, q->r, t->in_r, NULL);
            e += lerp(tV, q->r, t->in_r, NULL);
        }
        if (q->r < rcut_mag) {
            double X = dot(s + 3 * i, s + 3 * j) - s0[i] * s0[j];
            double J = lerp(tJ, q->r, t->in_r, NULL);
            dudr -= lerp(tJp, q->r, t->in_r, NULL) * X;
            e -= J * X;
            for (int c = 0; c < 3; ++c) {
                H[3 * i + c] += J * s[3 * j + c];
                H[3 * j + c] += J * s[3 * i + c];
            }
            sumJs[i] += J * s0[j];
            sumJs[j] += J * s0[i];
        }
        for (int c = 0; c < 3; ++c) {
            double fc = -dudr / q->r * q->d[c];
            F[3 * i + c] += fc;
            F[3 * j + c] -= fc;
        }
    }
    const double A = th[18], B = th[19], Cc = th[20], D = th[21];
    for (int i = 0; i < n; ++i) {
        double s2 = dot(s + 3 * i, s + 3 * i);
        e += A * s2 + B * pow(s2, 2) + Cc * pow(s2, 3) + D * pow(s2, 4);
        double kl = -(2e0 * A + 4e0 * B * s2 + 6e0 * Cc * pow(s2, 2) + 8e0 * D * pow(s2, 3));
        double hc = s0[i] > 0e0 ? -sumJs[i] / s0[i] : 0e0;
        for (int c = 0; c < 3; ++c) H[3 * i + c] += (kl + hc) * s[3 * i + c];
    }
    *E = e;

    /* derivatives, one parameter at a time */
    for (int k = 0; k < ORCM_NPARAM; ++k) {
        const int kind = kind_of(k);
        const double* b0 = k < 18 ? t->der[k][0] : NULL;
        const double* b1 = k < 18 ? t->der[k][1] : NULL;
        memset(drho, 0, sizeof(double) * n);
        memset(dF, 0, sizeof(double) * 3 * n);
        memset(dH, 0, sizeof(double) * 3 * n);
        double de = 0e0;
        if (kind == ORCM_DENS)
            for (size_t p = 0; p < np; ++p)
                if (P[p].r < rcut_pot) {
                    double db = lerp(b0, P[p].r, t->in_r, NULL);
                    drho[P[p].i] += db;
                    drho[P[p].j] += db;
                }
        for (int i = 0; i < n; ++i) {
            dfp[i] = sFp[i] * drho[i];
            de += sF[i] * drho[i];
            if (kind == ORCM_EMBED) {
                dfp[i] += lerp(b1, rho[i], t->in_rho, NULL);
                de += lerp(b0, rho[i], t->in_rho, NULL);
            }
        }
        double* dsumJs = sumJs; /* reused as d(sum J |s_j|) */
        memset(dsumJs, 0, sizeof(double) * n);
        for (size_t p = 0; p < np; ++p) {
            const pair_t* q = &P[p];
            const int i = q->i, j = q->j;
            double ddudr = 0e0;
            if (q->r < rcut_pot) {
                ddudr += (dfp[i] + dfp[j]) * lerp(tfp, q->r, t->in_r, NULL);
                if (kind == ORCM_DENS) ddudr += (fp[i] + fp[j]) * lerp(b1, q->r, t->in_r, NULL);
                if (kind == ORCM_PAIR) {
                    ddudr += lerp(b1, q->r, t->in_r, NULL);
                    de += lerp(b0, q->r, t->in_r, NULL);
                }
            }
            if (q->r < rcut_mag && kind == ORCM_EXCH) {
                double X = dot(s + 3 * i, s + 3 * j) - s0[i] * s0[j];
                double dJ = lerp(b0, q->r, t->in_r, NULL);
                ddudr -= lerp(b1, q->r, t->in_r, NULL) * X;
                de -= dJ * X;
                for (int c = 0; c < 3; ++c) {
                    dH[3 * i + c] += dJ * s[3 * j + c];
                    dH[3 * j + c] += dJ * s[3 * i + c];
                }
                dsumJs[i] += dJ * s0[j];
                dsumJs[j] += dJ * s0[i];
            }
            for (int c = 0; c < 3; ++c) {
                double fc = -ddudr / q->r * q->d[c];
                dF[3 * i + c] += fc;
                dF[3 * j + c] -= fc;
            }
        }
        for (int i = 0; i < n; ++i) {
            double s2 = dot(s + 3 * i, s + 3 * i);
            double dkl = 0e0;
            if (k == 18) { de += s2; dkl = -2e0; }
            if (k == 19) { de += pow(s2, 2); dkl = -4e0 * s2; }
            if (k == 20) { de += pow(s2, 3); dkl = -6e0 * pow(s2, 2); }
            if (k == 21) { de += pow(s2, 4); dkl = -8e0 * pow(s2, 3); }
            double dhc = s0[i] > 0e0 ? -dsumJs[i] / s0[i] : 0e0;
            for (int c = 0; c < 3; ++c) dH[3 * i + c] += (dkl + dhc) * s[3 * i + c];
        }
        double g = gE * de;
        for (int i = 0; i < 3 * n; ++i) {
            if (gF) g += gF[i] * dF[i];
            if (gH) g += gH[i] * dH[i];
        }
        grad[k] += g;
    }
    free(P); free(rho); free(fp); free(sF); free(sFp); free(s0); free(drho); free(dfp);
    free(dF); free(dH); free(sumJs);
    return 0;
}
