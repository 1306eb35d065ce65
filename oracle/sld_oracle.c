/*
 * sld_oracle.c - CPU restatement of the SPILADY SLDHL hot path (see
 * sld_oracle.h). TEST INFRASTRUCTURE ONLY: the checker, never the product.
 *
 * Every function cites the reference file:line it restates. Data structures:
 * atoms are indices 0..N-1 with the reference's per-cell doubly linked lists
 * kept as next/prev index arrays (struct.h:113-116), cells as in
 * cell_struct (struct.h:122-166). Single threaded; with the reference built
 * for one OpenMP thread the loop order, and so the floating-point order, is
 * the same.
 *
 * Parity status of each part is stated in the tests that pin it.
 */
#include "sld_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* spilady.h:53-62 */
#define BOLTZ 8.617343e-5
#define HBAR 6.58211899e-16
#define PI 3.141592653589793238462e0   /* Pi_num (spilady.h:61) */
#define EL_G 2.0023193043622e0
#define EVTOJ 1.602176487e-19

#define NINPUT 1000000
#define NTAB 16
enum { T_F, T_DF, T_SF, T_DSF, T_PR, T_DPR, T_J, T_DJ, T_LA, T_LB, T_LC, T_LD,
       T_DLA, T_DLB, T_DLC, T_DLD };

typedef struct { double x, y, z; } vec;

struct orc_ctx {
    sld_params prm;
    int natom, ncells, nx, ny, nz, ngroups, max_members;
    double d[6], inv[6];       /* xx yx yy zx zy zz */
    vec box_length;
    double box_volume;
    double gamma_L;            /* gamma_L_over_mass*atmass (read_variables.cpp:258-260) */
    double rcut_pot_sq, rcut_mag_sq, rcut_max_sq;
    /* tables (potential_CPU.cpp:61-157) */
    double rmax, rhomax, finput_over_rmax, finput_over_rhomax;
    double* tab[NTAB];
    /* atoms */
    vec *r, *p, *f, *s, *heff_h, *heff_l, *heff_hc;
    vec *hext, *fext;          /* external field (eV) and force (eV/A), extfield/extforce */
    double *s0, *rho, *pe, *ke, *me, *me0, *vir, *stress, *sum_jij_sj, *local_volume;
    double *sum_rij_m1, *sum_rij_m2;  /* localvol: sums of 1/r and 1/r^2 within rcut_vol (struct.h:105-108) */
    vec* m;                           /* magmom: the moments, the state's spin variable (struct.h) */
    int *next, *prev, *cell_old, *cell_new;
    /* cells */
    int *head, *tail, *count, *neigh, *type;
    int *group_size, *group_members;   /* [ngroups][max_members] */
    double *Te, *Ee, *Tl, *Gel, *Ges, *Ts_R, *Ts_L;
    double *sum_R_up, *sum_R_dn, *sum_L_up, *sum_L_dn;
    vec *ave_p, *ave_fluct;
    /* run state */
    double total_time, virial, initial_ave_energy;
    int rng_mode;
    uint32_t m_w, m_z;         /* MWC state (random_number_CPU.cpp:38-51) */
    int spare;
    double rand1, rand2;
    uint64_t noise_call;       /* counter-based noise: one id per noise pass */
    /* barostats (STRESS, PRESSURE) and the last check()'s values they act on */
    sld_barostat baro;
    double last_t_stress, last_t_pressure;
    double last_stress[6], last_pressure;
    /* quantum thermostat (quantum_noise_CPU.cpp): filter, per-DOF history
     * [3N][nf2] oldest first, current noise and call count per DOF */
    int qn_nf2;
    double *qH, *qmem, *qval;
    int* qcount;
};

/* ------------------------------------------------------------------------ */
/* vec_utils_CPU.cpp:33-148                                                  */
static inline vec v_add(vec a, vec b) { vec c = {a.x + b.x, a.y + b.y, a.z + b.z}; return c; }
static inline vec v_sub(vec a, vec b) { vec c = {a.x - b.x, a.y - b.y, a.z - b.z}; return c; }
static inline vec v_cross(vec a, vec b) {
    vec c = {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
    return c;
}
static inline vec v_times(double a, vec b) { vec c = {a * b.x, a * b.y, a * b.z}; return c; }
static inline vec v_divide(vec a, double b) { vec c = {a.x / b, a.y / b, a.z / b}; return c; }
static inline double v_dot(vec a, vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double v_sq(vec a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double v_length(vec a) { return sqrt(v_sq(a)); }
static inline vec v_zero(void) { vec c = {0e0, 0e0, 0e0}; return c; }

/* the start of an atom's spin field sum: Hext with extfield, else zero
 * (calculate_spin_CPU.cpp:35-39, calculate_temperature_CPU.cpp:71-75,
 * check_temperature_CPU.cpp:59-63) */
static inline vec hext_of(const orc_ctx* o, int a) {
    return o->prm.extfield ? o->hext[a] : v_zero();
}

/* the force of the momentum update: with extforce, f += fext in place first
 * (core_dp_CPU.cpp:69-71,130-132) */
static inline vec dp_force(orc_ctx* o, int a) {
    if (o->prm.extforce) o->f[a] = v_add(o->f[a], o->fext[a]);
    return o->f[a];
}

/* ------------------------------------------------------------------------ */
/* DD05 potential (DD05.cpp:6-164) and Fe exchange/Landau (JijFe.cpp:5-43)   */
static double bigf_gen(double rho) {
    double aa = 4.100199340884814;
    double bb = 1.565647547483517;
    if (rho >= 1e0) return -aa * sqrt(rho);
    return -aa * sqrt(rho) - bb / log(2e0) * (1e0 - sqrt(rho)) * log(2e0 - rho);
}

static const double SF_A[6] = {0.9332056681088162, -1.162558782567700, -0.3502026949249225,
                               0.4287820835430028, 4.907925057809273, -5.307049068415304};
static const double SF_R[6] = {3.0, 2.8666666666666666, 2.7333333333333333, 2.6, 2.4, 2.3};

static double smallf_gen(double rij) {
    double f = 0e0;
    for (int i = 0; i < 6; ++i)
        if (rij < SF_R[i]) f += SF_A[i] * pow(SF_R[i] - rij, 3);
    return f;
}

static double pairZBL(double rij) {
    double bb0 = 0.1818e0, bb1 = 0.5099e0, bb2 = 0.2802e0, bb3 = 0.02817e0;
    double nzz = 26e0;
    double abohr = 0.52917720859e0;
    double as = 0.88535e0 * abohr / (2e0 * pow(nzz, 0.23));
    double xx = rij / as;
    double pair = bb0 * exp(-3.2 * xx) + bb1 * exp(-0.9423 * xx) + bb2 * exp(-0.4029 * xx) +
                  bb3 * exp(-0.2016 * xx);
    pair = nzz * nzz * 14.3992 * pair / rij;
    return pair;
}

static const double PV[8] = {-0.1960674387419232, 0.3687525935422963, -1.505333614924853,
                             4.948907078156191,   -4.894613262753399, 3.468897724782442,
                             -1.792218099820337,  80.22069592246987};
static const double PR[8] = {4.1, 3.8, 3.5, 3.2, 2.9, 2.6, 2.4, 2.3};

/* DD05.cpp:88-158: quintic through the ZBL end (x1 = 1.8) and the spline
 * end (x2 = 2.0) matching value, first and second derivative at both ends;
 * ends = {y1, Dy1, D2y1, y2, Dy2, D2y2}, a[k] the coefficient of r^k */
void orc_bridge_coeffs(const double ends[6], double a[6]) {
    double x1 = 1.8, x2 = 2.0;
    double y1 = ends[0], Dy1 = ends[1], D2y1 = ends[2], y2 = ends[3], Dy2 = ends[4], D2y2 = ends[5];
    a[5] = (-6e0*Dy1*x1-6e0*Dy2*x1+D2y1*pow(x1,2)-D2y2*pow(x1,2)
                   +6e0*Dy1*x2+6e0*Dy2*x2-2e0*D2y1*x1*x2+2e0*D2y2*x1*x2
                   +D2y1*pow(x2,2)-D2y2*pow(x2,2)+12*y1-12*y2)
                   /(2e0*pow((x1-x2),5));
    a[4] = (-2e0*D2y1*pow(x1,3)+3e0*D2y2*pow(x1,3)+D2y1*pow(x1,2)*x2
                   -4e0*D2y2*pow(x1,2)*x2+4*D2y1*x1*pow(x2,2)-D2y2*x1*pow(x2,2)
                   -3e0*D2y1*pow(x2,3)+2e0*D2y2*pow(x2,3)
                   +2e0*Dy1*(7e0*pow(x1,2)+x1*x2-8e0*pow(x2,2))
                   +2e0*Dy2*(8e0*pow(x1,2)-x1*x2-7e0*pow(x2,2))
                   -30e0*x1*y1-30e0*x2*y1+30e0*x1*y2+30e0*x2*y2)
                   /(2e0*pow((x1-x2),5));
    a[3] = -(-D2y1*pow(x1,4)+3e0*D2y2*pow(x1,4)-4e0*D2y1*pow(x1,3)*x2
                   +8e0*D2y1*pow(x1,2)*pow(x2,2)-8e0*D2y2*pow(x1,2)*pow(x2,2)
                   +4e0*D2y2*x1*pow(x2,3)-3e0*D2y1*pow(x2,4)+D2y2*pow(x2,4)
                   +4e0*Dy1*(2e0*pow(x1,3)+8e0*pow(x1,2)*x2-7e0*x1*pow(x2,2)-3e0*pow(x2,3))
                   +4e0*Dy2*(3e0*pow(x1,3)+7e0*pow(x1,2)*x2-8e0*x1*pow(x2,2)-2e0*pow(x2,3))
                   -20e0*pow(x1,2)*y1-80e0*x1*x2*y1-20e0*pow(x2,2)*y1
                   +20e0*pow(x1,2)*y2+80e0*x1*x2*y2+20e0*pow(x2,2)*y2)
                   /(2e0*pow((x1-x2),5));
    a[2] = (D2y2*x1*pow((x1-x2),2)*(pow(x1,2)+6e0*x1*x2+3e0*pow(x2,2))
                   -x2*(3e0*D2y1*pow(x1,4)-8e0*D2y1*pow(x1,2)*pow(x2,2)
                   +4e0*D2y1*x1*pow(x2,3)+D2y1*pow(x2,4)
                   -12e0*Dy1*x1*(2e0*pow(x1,2)+x1*x2-3e0*pow(x2,2))
                   +12e0*Dy2*x1*(-3e0*pow(x1,2)+x1*x2+2e0*pow(x2,2))
                   +60e0*pow(x1,2)*y1+60e0*x1*x2*y1
                   -60e0*pow(x1,2)*y2-60e0*x1*x2*y2))
                   /(2e0*pow((x1-x2),5));
    a[1] = (2e0*Dy2*pow(x1,2)*(pow(x1,3)-5e0*pow(x1,2)*x2-8e0*x1*pow(x2,2)+12e0*pow(x2,3))
                  +x2*(-D2y2*pow(x1,2)*pow((x1-x2),2)*(2e0*x1+3e0*x2)
                  +x2*(-2e0*Dy1*(12e0*pow(x1,3)-8e0*pow(x1,2)*x2-5e0*x1*pow(x2,2)+pow(x2,3))
                  +x1*(D2y1*pow((x1-x2),2)*(3e0*x1+2e0*x2)+60e0*x1*(y1-y2)))))
                  /(2e0*pow((x1-x2),5));
    a[0] = (D2y2*pow(x1,3)*pow((x1-x2),2)*pow(x2,2)+8e0*Dy1*pow(x1,3)*pow(x2,3)
                   -D2y1*pow(x1,4)*pow(x2,3)-10e0*Dy1*pow(x1,2)*pow(x2,4)
                   +2e0*D2y1*pow(x1,3)*pow(x2,4)+2e0*Dy1*x1*pow(x2,5)-D2y1*pow(x1,2)*pow(x2,5)
                   -2e0*Dy2*pow(x1,3)*x2*(pow(x1,2)-5e0*x1*x2+4e0*pow(x2,2))
                   -20e0*pow(x1,2)*pow(x2,3)*y1+10e0*x1*pow(x2,4)*y1-2e0*pow(x2,5)*y1
                   +2e0*pow(x1,5)*y2-10e0*pow(x1,4)*x2*y2+20e0*pow(x1,3)*pow(x2,2)*y2)
                   /(2e0*pow((x1-x2),5));
}

/* DD05.cpp:64-164 with the knot coefficients V[8]: ZBL below 1.8 A, the
 * quintic bridge to 2.0 A, cubic knots above */
double orc_pair_v(double rij, const double V[8]) {
    const int N = 8;
    double pair = 0e0;
    double x1 = 1.8, x2 = 2.0, offset = 0e0;
    if (rij <= x1) {
        pair = pairZBL(rij) + offset;
    } else if (rij <= x2) {
        double e[6], a[6];
        e[0] = pairZBL(x1) + offset;
        e[1] = (pairZBL(x1 + 1e-7) - pairZBL(x1 - 1e-7)) / 2e-7;
        e[2] = (pairZBL(x1 + 1e-7) + pairZBL(x1 - 1e-7) - 2e0 * pairZBL(x1)) / pow(1e-7, 2);
        double y2 = 0e0, Dy2 = 0e0, D2y2 = 0e0;
        for (int i = 0; i < N; ++i) if (x2 < PR[i]) y2 += V[i] * pow((PR[i] - x2), 3);
        for (int i = 0; i < N; ++i) if (x2 < PR[i]) Dy2 -= 3e0 * V[i] * pow((PR[i] - x2), 2);
        for (int i = 0; i < N; ++i) if (x2 < PR[i]) D2y2 += 6e0 * V[i] * (PR[i] - x2);
        e[3] = y2; e[4] = Dy2; e[5] = D2y2;
        orc_bridge_coeffs(e, a);
        pair = a[5]*pow(rij,5)+a[4]*pow(rij,4)+a[3]*pow(rij,3)+a[2]*pow(rij,2)+a[1]*rij+a[0];
    } else {
        for (int i = 0; i < N; ++i)
            if (rij < PR[i]) pair += V[i] * pow((PR[i] - rij), 3);
    }
    return pair;
}

static double pair_gen(double rij) { return orc_pair_v(rij, PV); }

static double Jij_gen(double rij) {
    double A = 0.870, rc = 3.75;
    double J = 0e0;
    if (rij < rc) J = A * pow(1.0 - rij / rc, 3);
    return J;
}
static double LandauA_gen(double rho) { (void)rho; return -0.744824; }
static double LandauB_gen(double rho) { (void)rho; return 0.345295; }
static double LandauC_gen(double rho) { (void)rho; return -0.00790205; }
static double LandauD_gen(double rho) { (void)rho; return 0e0; }

/* potential_CPU.cpp:159-170: 6-point central difference, dx = x/1000 */
static double first_derivative(double (*fn)(double), double x) {
    double dx = x / 1000e0;
    double vp = fn(x + dx), vm = fn(x - dx);
    double v2p = fn(x + 2 * dx), v2m = fn(x - 2 * dx);
    double v3p = fn(x + 3 * dx), v3m = fn(x - 3 * dx);
    return (45e0 * (vp - vm) - 9e0 * (v2p - v2m) + (v3p - v3m)) / (60e0 * dx);
}

double orc_analytic(int which, double x) {
    switch (which) {
        case T_F: return bigf_gen(x);
        case T_DF: return first_derivative(bigf_gen, x);
        case T_SF: return smallf_gen(x);
        case T_DSF: return first_derivative(smallf_gen, x);
        case T_PR: return pair_gen(x);
        case T_DPR: return first_derivative(pair_gen, x);
        case T_J: return Jij_gen(x);
        case T_DJ: return first_derivative(Jij_gen, x);
        case T_LA: return LandauA_gen(x);
        case T_LB: return LandauB_gen(x);
        case T_LC: return LandauC_gen(x);
        case T_LD: return LandauD_gen(x);
        case T_DLA: return first_derivative(LandauA_gen, x);
        case T_DLB: return first_derivative(LandauB_gen, x);
        case T_DLC: return first_derivative(LandauC_gen, x);
        case T_DLD: return first_derivative(LandauD_gen, x);
    }
    return NAN;
}

/* potential_CPU.cpp:61-157 */
static void potential_table(orc_ctx* o) {
    o->rmax = o->prm.rcut_max;
    o->rhomax = 2e0 * smallf_gen(1e0);
    if (o->rhomax < 1000e0) o->rhomax = 1000e0;
    double finput = (double)NINPUT;
    o->finput_over_rmax = finput / o->rmax;
    o->finput_over_rhomax = finput / o->rhomax;
    for (int t = 0; t < NTAB; ++t) o->tab[t] = (double*)malloc(sizeof(double) * NINPUT);
    for (int n = 0; n < NINPUT; ++n) {
        double rij = o->rmax / (double)NINPUT * (double)(n + 1);
        double rho = o->rhomax / (double)NINPUT * (double)(n + 1);
        o->tab[T_F][n] = bigf_gen(rho);
        o->tab[T_DF][n] = first_derivative(bigf_gen, rho);
        o->tab[T_SF][n] = smallf_gen(rij);
        o->tab[T_DSF][n] = first_derivative(smallf_gen, rij);
        o->tab[T_PR][n] = pair_gen(rij);
        o->tab[T_DPR][n] = first_derivative(pair_gen, rij);
        o->tab[T_J][n] = Jij_gen(rij);
        o->tab[T_DJ][n] = first_derivative(Jij_gen, rij);
        o->tab[T_LA][n] = LandauA_gen(rho);
        o->tab[T_LB][n] = LandauB_gen(rho);
        o->tab[T_LC][n] = LandauC_gen(rho);
        o->tab[T_LD][n] = LandauD_gen(rho);
        o->tab[T_DLA][n] = first_derivative(LandauA_gen, rho);
        o->tab[T_DLB][n] = first_derivative(LandauB_gen, rho);
        o->tab[T_DLC][n] = first_derivative(LandauC_gen, rho);
        o->tab[T_DLD][n] = first_derivative(LandauD_gen, rho);
        if (o->prm.magmom) { /* potential_CPU.cpp:125-160 */
            o->tab[T_J][n] *= pow(EL_G, 2);
            o->tab[T_DJ][n] *= pow(EL_G, 2);
            o->tab[T_LA][n] *= pow(EL_G, 2);
            o->tab[T_LB][n] *= pow(EL_G, 4);
            o->tab[T_LC][n] *= pow(EL_G, 6);
            o->tab[T_LD][n] *= pow(EL_G, 8);
            o->tab[T_DLA][n] *= pow(EL_G, 2);
            o->tab[T_DLB][n] *= pow(EL_G, 4);
            o->tab[T_DLC][n] *= pow(EL_G, 6);
            o->tab[T_DLD][n] *= pow(EL_G, 8);
        }
    }
}

/* potential_CPU.cpp:172-179. n = 0 would read entry -1 (undefined in the
 * reference); the restatement clamps it to 1, which only matters for x below
 * one table step. */
static inline double interp(double x, double in_over_max, const double* t) {
    double fn = x * in_over_max;
    int n = (int)fn;
    if (n < 1) n = 1;
    double ratio = (double)(n + 1) - fn;
    return t[n - 1] * ratio + t[n] * (1e0 - ratio);
}
#define TR(o, w, x) interp((x), (o)->finput_over_rmax, (o)->tab[w])
#define TRHO(o, w, x) interp((x), (o)->finput_over_rhomax, (o)->tab[w])

double orc_table(const orc_ctx* o, int which, double x) {
    if (which < 0 || which >= NTAB) return NAN;
    int rho_kind = (which == T_F || which == T_DF || which >= T_LA);
    return interp(x, rho_kind ? o->finput_over_rhomax : o->finput_over_rmax, o->tab[which]);
}
double orc_table_raw(const orc_ctx* o, int which, int n) {
    if (which < 0 || which >= NTAB || n < 0 || n >= NINPUT) return NAN;
    return o->tab[which][n];
}

/* heatcapacity_CPU.cpp:34-57 */
static double Ce(double Te) { double a = 3e0, b = 2e-4 / BOLTZ; return a * tanh(b * Te); }
static double Te_to_Ee(double Te) { double a = 3e0, b = 2e-4 / BOLTZ; return a / b * log(cosh(b * Te)); }
static double Ee_to_Te(double Ee) { double a = 3e0, b = 2e-4 / BOLTZ; return acosh(exp(b / a * Ee)) / b; }

/* ------------------------------------------------------------------------ */
/* noise                                                                     */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), multipliers and Weyl
 * constants as published there. */
void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Counter-based noise of one atom in one noise pass: key = (seed low,
 * seed high ^ replica), counter = (atom, pass low, pass high, 'SPLD').
 * Four 32-bit words become uniforms in (0,1) and two Box-Muller pairs, of
 * which the first three normals are used. */
void orc_normal3(uint64_t seed, int replica, int atom, uint64_t call, double out[3]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32) ^ (uint32_t)replica};
    uint32_t ctr[4] = {(uint32_t)atom, (uint32_t)call, (uint32_t)(call >> 32), 0x53504C44u};
    uint32_t w[4];
    orc_philox4x32(ctr, key, w);
    const double inv32 = 2.3283064365386963e-10; /* 2^-32 */
    double u0 = ((double)w[0] + 0.5) * inv32;
    double u1 = ((double)w[1] + 0.5) * inv32;
    double u2 = ((double)w[2] + 0.5) * inv32;
    double u3 = ((double)w[3] + 0.5) * inv32;
    const double two_pi = 6.283185307179586;
    double a = sqrt(-2e0 * log(u0));
    double b = sqrt(-2e0 * log(u2));
    out[0] = a * cos(two_pi * u1);
    out[1] = a * sin(two_pi * u1);
    out[2] = b * cos(two_pi * u3);
}

/* random_number_CPU.cpp:53-83, one stream (thread 0) */
static double mwc_uniform(orc_ctx* o) {
    uint32_t m_z = o->m_z, m_w = o->m_w;
    m_z = 36969 * (m_z & 65535) + (m_z >> 16);
    m_w = 18000 * (m_w & 65535) + (m_w >> 16);
    uint32_t un = (m_z << 16) + m_w;
    o->m_z = m_z; o->m_w = m_w;
    return (un + 1.0) * 2.328306435454494e-10;
}
static double mwc_normal(orc_ctx* o) {
    if (o->spare == 1) {
        o->spare = 0;
        return sqrt(o->rand1) * sin(o->rand2);
    }
    o->spare = 1;
    o->rand1 = mwc_uniform(o);
    o->rand1 = -2e0 * log(o->rand1);
    o->rand2 = mwc_uniform(o) * 6.283185307179586;
    return sqrt(o->rand1) * cos(o->rand2);
}

/* The counter-based stream of the quantum noise history: element e of an
 * atom's three histories is normal3 at call QN_CALL0 + e, a range no lattice
 * or spin pass reaches. */
#define QN_CALL0 (1ull << 62)

static inline void noise3(orc_ctx* o, int atom, double out[3]);

/* initial_quantum_noise (quantum_noise_CPU.cpp:83-130): the filter H is the
 * cosine transform of H~(w) = sqrt(hbar|w| (1/2 + 1/expm1(hbar|w|/T)))
 * * (w h/2) / sin(w h/2) on 2 Nf frequencies up to pi/h, h = M dt / 2
 * (H~(0) = sqrt(T)); every history starts with fresh normals. */
static void quantum_init(orc_ctx* o) {
    const int Nf = o->prm.nfrequency_quantum, Nf2 = 2 * Nf, n = 3 * o->natom;
    const double T = o->prm.temperature;
    const double h = o->prm.msteps_quantum * o->prm.step / 2e0;
    const double max_frequency = PI / h;
    const double delta_frequency = max_frequency / (double)Nf;
    o->qn_nf2 = Nf2;
    o->qH = (double*)calloc((size_t)Nf2, sizeof(double));
    double* Ht = (double*)calloc((size_t)Nf2, sizeof(double));
    for (int i = 0; i < Nf2; ++i) {
        int k = i - Nf;
        double frequency = delta_frequency * (double)k;
        Ht[i] = sqrt(HBAR * fabs(frequency) * (0.5e0 + 1e0 / expm1(HBAR * fabs(frequency) / T)));
        Ht[i] *= (frequency * h / 2e0) / sin(frequency * h / 2e0);
        if (k == 0) Ht[i] = sqrt(T);
    }
    for (int i = 0; i < Nf2; ++i) {
        o->qH[i] = 0e0;
        for (int j = 0; j < Nf2; ++j) {
            int k = j - Nf;
            o->qH[i] += Ht[j] * cos(PI / (double)Nf * (double)k * (double)(i - Nf));
        }
        o->qH[i] /= (double)Nf2;
    }
    free(Ht);
    o->qmem = (double*)calloc((size_t)n * Nf2, sizeof(double));
    o->qval = (double*)calloc((size_t)n, sizeof(double));
    o->qcount = (int*)calloc((size_t)n, sizeof(int));
    if (o->rng_mode == ORC_RNG_MWC) {
        for (size_t i = 0; i < (size_t)n * Nf2; ++i) o->qmem[i] = mwc_normal(o);
    } else {
        for (int a = 0; a < o->natom; ++a)
            for (int j = 0; j < Nf2; ++j) {
                double z[3];
                orc_normal3(o->prm.random_seed, o->prm.replica, a, QN_CALL0 + (uint64_t)j, z);
                for (int k = 0; k < 3; ++k) o->qmem[(size_t)(3 * a + k) * Nf2 + j] = z[k];
            }
    }
}

/* quantum_noise (quantum_noise_CPU.cpp:60-81): every M-th call of a degree
 * of freedom, its noise becomes the filter applied to its history (newest
 * against H[0]), the history shifts by one and takes a fresh normal. */
static double quantum_noise(orc_ctx* o, int atom, int k) {
    const int Nf2 = o->qn_nf2, n = 3 * atom + k;
    double* mem = o->qmem + (size_t)n * Nf2;
    if (o->qcount[n] % o->prm.msteps_quantum == 0) {
        double noise = 0e0;
        for (int i = 0; i < Nf2; ++i) noise += o->qH[i] * mem[Nf2 - i - 1];
        for (int i = 0; i < Nf2 - 1; ++i) mem[i] = mem[i + 1];
        if (o->rng_mode == ORC_RNG_MWC) {
            mem[Nf2 - 1] = mwc_normal(o);
        } else {
            /* the r-th refresh appends history element Nf2 + r */
            uint64_t e = (uint64_t)Nf2 + (uint64_t)(o->qcount[n] / o->prm.msteps_quantum);
            double z[3];
            orc_normal3(o->prm.random_seed, o->prm.replica, atom, QN_CALL0 + e, z);
            mem[Nf2 - 1] = z[k];
        }
        o->qval[n] = noise;
    }
    ++o->qcount[n];
    return o->qval[n];
}

/* the random force of one atom in core_dp_A (core_dp_CPU.cpp:134-155):
 * white noise of variance 2 T gamma / dt, or with quantumnoise the coloured
 * noise scaled by sqrt(2 gamma / h), h = M dt */
static vec fluct_force(orc_ctx* o, int atom, double T, double gl, double dt) {
    double nz[3], len;
    if (o->prm.quantumnoise) {
        double h = dt * o->prm.msteps_quantum;
        len = sqrt(2e0 * gl / h);
        nz[0] = quantum_noise(o, atom, 0);
        nz[1] = quantum_noise(o, atom, 1);
        nz[2] = quantum_noise(o, atom, 2);
    } else {
        len = sqrt(2e0 * T * gl / dt);
        noise3(o, atom, nz);
    }
    vec ff = {nz[0], nz[1], nz[2]};
    return v_times(len, ff);
}

static inline void noise3(orc_ctx* o, int atom, double out[3]) {
    if (o->rng_mode == ORC_RNG_MWC) {
        out[0] = mwc_normal(o); out[1] = mwc_normal(o); out[2] = mwc_normal(o);
    } else {
        orc_normal3(o->prm.random_seed, o->prm.replica, atom, o->noise_call, out);
    }
}

/* ------------------------------------------------------------------------ */
/* geometry                                                                  */

/* vec_utils_CPU.cpp:135-148 */
static void inverse_box(const double* d, double* inv) {
    inv[0] = 1e0 / d[0];
    inv[1] = -d[1] / (d[0] * d[2]);
    inv[2] = 1e0 / d[2];
    inv[3] = (d[1] * d[4] - d[2] * d[3]) / (d[0] * d[2] * d[5]);
    inv[4] = -d[4] / (d[2] * d[5]);
    inv[5] = 1e0 / d[5];
}

/* periodic_CPU.cpp:31-50, including its treatment of q in (-1, 0) */
static void periodic(const orc_ctx* o, vec* r) {
    const double *d = o->d, *I = o->inv;
    vec q;
    q.x = I[0] * r->x + I[1] * r->y + I[3] * r->z;
    q.y = I[2] * r->y + I[4] * r->z;
    q.z = I[5] * r->z;
    int index = 0;
    if (q.x < 0e0 || q.x >= 1e0) { q.x -= (int)q.x; ++index; }
    if (q.y < 0e0 || q.y >= 1e0) { q.y -= (int)q.y; ++index; }
    if (q.z < 0e0 || q.z >= 1e0) { q.z -= (int)q.z; ++index; }
    if (index > 0) {
        r->x = d[0] * q.x + d[1] * q.y + d[3] * q.z;
        r->y = d[2] * q.y + d[4] * q.z;
        r->z = d[5] * q.z;
    }
}

/* find_image_CPU.cpp:31-53 */
static inline void find_image(const orc_ctx* o, vec* rij) {
    const double *d = o->d, *I = o->inv;
    vec q;
    q.x = I[0] * rij->x + I[1] * rij->y + I[3] * rij->z;
    q.y = I[2] * rij->y + I[4] * rij->z;
    q.z = I[5] * rij->z;
    int index = 0;
    if (q.x < -0.5e0) { q.x += 1e0; ++index; }
    if (q.x >= 0.5e0) { q.x -= 1e0; ++index; }
    if (q.y < -0.5e0) { q.y += 1e0; ++index; }
    if (q.y >= 0.5e0) { q.y -= 1e0; ++index; }
    if (q.z < -0.5e0) { q.z += 1e0; ++index; }
    if (q.z >= 0.5e0) { q.z -= 1e0; ++index; }
    if (index > 0) {
        rij->x = d[0] * q.x + d[1] * q.y + d[3] * q.z;
        rij->y = d[2] * q.y + d[4] * q.z;
        rij->z = d[5] * q.z;
    }
}

/* map_cells.cpp:41-121 */
static void map_cells(orc_ctx* o) {
    const double* d = o->d;
    double mlc = o->prm.min_length_link_cell;
    double angle1 = sin(atan2(d[2], d[1]));
    double angle2 = sin(atan2(d[5], d[3]));
    double angle3 = sin(atan2(d[5], d[4]));
    double ex = fmax(mlc / angle1, mlc / angle2);
    double ey = fmax(mlc, mlc / angle3);
    double ez = mlc;
    o->nx = (int)(d[0] / ex);
    o->ny = (int)(d[2] / ey);
    o->nz = (int)(d[5] / ez);
    o->ncells = o->nx * o->ny * o->nz;
    int nx = o->nx, ny = o->ny, nz = o->nz, nd = nx, nh = nx * ny;
    o->neigh = (int*)malloc(sizeof(int) * 26 * o->ncells);
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                int c = ix + iy * nd + iz * nh;
                int xp = (ix + 1) % nx, yp = (iy + 1) % ny, zp = (iz + 1) % nz;
                int xm = (ix - 1 + nx) % nx, ym = (iy - 1 + ny) % ny, zm = (iz - 1 + nz) % nz;
                int* n = o->neigh + 26 * c;
                n[0] = xp + iy * nd + iz * nh;   n[1] = xp + yp * nd + iz * nh;
                n[2] = ix + yp * nd + iz * nh;   n[3] = xm + yp * nd + iz * nh;
                n[4] = xp + iy * nd + zm * nh;   n[5] = xp + yp * nd + zm * nh;
                n[6] = ix + yp * nd + zm * nh;   n[7] = xm + yp * nd + zm * nh;
                n[8] = xp + iy * nd + zp * nh;   n[9] = xp + yp * nd + zp * nh;
                n[10] = ix + yp * nd + zp * nh;  n[11] = xm + yp * nd + zp * nh;
                n[12] = ix + iy * nd + zp * nh;  n[13] = xm + iy * nd + iz * nh;
                n[14] = ix + ym * nd + iz * nh;  n[15] = ix + iy * nd + zm * nh;
                n[16] = xp + ym * nd + iz * nh;  n[17] = xm + ym * nd + iz * nh;
                n[18] = xm + iy * nd + zp * nh;  n[19] = xm + iy * nd + zm * nh;
                n[20] = ix + ym * nd + zp * nh;  n[21] = ix + ym * nd + zm * nh;
                n[22] = xp + ym * nd + zp * nh;  n[23] = xp + ym * nd + zm * nh;
                n[24] = xm + ym * nd + zp * nh;  n[25] = xm + ym * nd + zm * nh;
            }
}

static inline int bin_atom(const orc_ctx* o, vec r, int clamp) {
    const double* I = o->inv;
    vec q;
    q.x = I[0] * r.x + I[1] * r.y + I[3] * r.z;
    q.y = I[2] * r.y + I[4] * r.z;
    q.z = I[5] * r.z;
    int lx = (int)(q.x * o->nx), ly = (int)(q.y * o->ny), lz = (int)(q.z * o->nz);
    if (clamp) {
        if (lx < 0) lx = 0;
        if (ly < 0) ly = 0;
        if (lz < 0) lz = 0;
        if (lx >= o->nx) lx = o->nx - 1;
        if (ly >= o->ny) ly = o->ny - 1;
        if (lz >= o->nz) lz = o->nz - 1;
    }
    return lx + ly * o->nx + lz * o->nx * o->ny;
}

static inline void push_head(orc_ctx* o, int i, int c) {
    o->prev[i] = -1;
    o->next[i] = o->head[c];
    o->head[c] = i;
    if (o->next[i] == -1) o->tail[c] = i;
    else o->prev[o->next[i]] = i;
}

/* links_CPU.cpp:33-83. The reference does not clamp here; out-of-range
 * indices are undefined there, so the restatement clamps as links() does. */
static void initial_links(orc_ctx* o) {
    for (int c = 0; c < o->ncells; ++c) { o->head[c] = -1; o->tail[c] = -1; o->count[c] = 0; }
    for (int i = 0; i < o->natom; ++i) {
        o->next[i] = -1; o->prev[i] = -1;
        o->cell_old[i] = -1; o->cell_new[i] = -2;
        periodic(o, &o->r[i]);
    }
    for (int i = 0; i < o->natom; ++i) {
        int c = bin_atom(o, o->r[i], 1);
        o->cell_new[i] = c;
        push_head(o, i, c);
        ++o->count[c];
    }
}

/* links_CPU.cpp:85-177: movers are detached and pushed at the head of their
 * new cell in ascending atom index (the per-thread chunks of the static
 * schedule are ascending and are replayed in thread order). */
void orc_links(orc_ctx* o) {
    int nrel = 0;
    for (int i = 0; i < o->natom; ++i) {
        o->cell_old[i] = o->cell_new[i];
        int c = bin_atom(o, o->r[i], 1);
        o->cell_new[i] = c;
        if (c != o->cell_old[i]) ++nrel;
    }
    if (nrel == 0) return;
    for (int i = 0; i < o->natom; ++i) {
        if (o->cell_new[i] == o->cell_old[i]) continue;
        int co = o->cell_old[i], cn = o->cell_new[i];
        if (o->prev[i] == -1) o->head[co] = o->next[i];
        else o->next[o->prev[i]] = o->next[i];
        if (o->next[i] == -1) o->tail[co] = o->prev[i];
        else o->prev[o->next[i]] = o->prev[i];
        push_head(o, i, cn);
        --o->count[co];
        ++o->count[cn];
    }
}

/* allocate_cells_CPU.cpp:33-125: greedy colouring in cell index order */
static void allocate_cells(orc_ctx* o) {
    int C = o->ncells;
    int* type = o->type;
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < C; ++i) type[i] = 0;
        int ng = 0, left = C;
        if (pass == 0) o->max_members = 0;
        while (1) {
            int nm = 0;
            for (int i = 0; i < C; ++i) {
                if (type[i] == 0) {
                    if (pass == 1) o->group_members[ng * o->max_members + nm] = i;
                    type[i] = 1;
                    ++nm;
                    --left;
                    for (int j = 0; j < 26; ++j) {
                        int nb = o->neigh[26 * i + j];
                        if (type[nb] != 1) type[nb] = -1;
                    }
                }
            }
            if (pass == 1) o->group_size[ng] = nm;
            ++ng;
            for (int i = 0; i < C; ++i) if (type[i] == -1) type[i] = 0;
            if (pass == 0 && nm > o->max_members) o->max_members = nm;
            if (left == 0) break;
        }
        o->ngroups = ng;
        if (pass == 0) {
            o->group_size = (int*)calloc(ng, sizeof(int));
            o->group_members = (int*)malloc(sizeof(int) * (size_t)ng * o->max_members);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* density, force and energy                                                 */

/* calculate_rho_CPU.cpp:33-66: half shell (cells 0..12, own cell up to self) */
static void embedded_rho(orc_ctx* o, int a) {
    int cc = o->cell_new[a];
    for (int i = 0; i <= 13; ++i) {
        int wc = (i == 13) ? cc : o->neigh[26 * cc + i];
        for (int w = o->head[wc]; w != -1; w = o->next[w]) {
            if (w == a && i == 13) break;
            vec rij = v_sub(o->r[a], o->r[w]);
            find_image(o, &rij);
            double rij0 = v_length(rij);
            if (rij0 < o->prm.rcut_pot && a != w) {
                double sf = TR(o, T_SF, rij0);
                o->rho[a] += sf;
                o->rho[w] += sf;
            }
        }
    }
}

/* calculate_rho_CPU.cpp:68-92 */
void orc_calculate_rho(orc_ctx* o) {
    for (int i = 0; i < o->natom; ++i) o->rho[i] = 0e0;
    for (int g = 0; g < o->ngroups; ++g)
        for (int j = 0; j < o->group_size[g]; ++j) {
            int c = o->group_members[g * o->max_members + j];
            for (int a = o->head[c]; a != -1; a = o->next[a]) embedded_rho(o, a);
        }
}

/* calculate_force_energy_CPU.cpp:33-216 (SLDHL; localvol's sums at :80-90) */
static void inner_loop(orc_ctx* o, int a) {
    double si_sq = v_sq(o->s[a]);
    int cc = o->cell_new[a];
    double* st_a = o->stress + 6 * a;
    for (int i = 0; i <= 13; ++i) {
        int wc = (i == 13) ? cc : o->neigh[26 * cc + i];
        for (int w = o->head[wc]; w != -1; w = o->next[w]) {
            if (w == a && i == 13) break;
            vec rij = v_sub(o->r[a], o->r[w]);
            find_image(o, &rij);
            double rsq = v_sq(rij);
            double dudr = 0e0;
            double dudr_spin = 0e0;
            double sj_sq = v_sq(o->s[w]);
            if (rsq < o->rcut_max_sq && a != w) {
                double rij0 = sqrt(rsq);
                if (o->prm.localvol && rij0 < o->prm.rcut_vol) {
                    o->sum_rij_m1[a] += 1e0 / rij0;
                    o->sum_rij_m2[a] += 1e0 / rsq;
                    o->sum_rij_m1[w] += 1e0 / rij0;
                    o->sum_rij_m2[w] += 1e0 / rsq;
                }
                if (rij0 < o->prm.rcut_pot) {
                    double dsf = TR(o, T_DSF, rij0);
                    double ri = o->rho[a], rj = o->rho[w];
                    dudr = (TRHO(o, T_DF, ri) + TRHO(o, T_DF, rj)) * dsf + TR(o, T_DPR, rij0);
                    dudr += (TRHO(o, T_DLA, ri) * si_sq + TRHO(o, T_DLB, ri) * pow(si_sq, 2) +
                             TRHO(o, T_DLC, ri) * pow(si_sq, 3) + TRHO(o, T_DLD, ri) * pow(si_sq, 4)) * dsf;
                    dudr += (TRHO(o, T_DLA, rj) * sj_sq + TRHO(o, T_DLB, rj) * pow(sj_sq, 2) +
                             TRHO(o, T_DLC, rj) * pow(sj_sq, 3) + TRHO(o, T_DLD, rj) * pow(sj_sq, 4)) * dsf;
                    double pair_enr = TR(o, T_PR, rij0);
                    o->pe[a] += 0.5e0 * pair_enr;
                    o->pe[w] += 0.5e0 * pair_enr;
                }
                if (rij0 < o->prm.rcut_mag) {
                    double si_dot_sj = v_dot(o->s[a], o->s[w]);
                    double si_times_sj = sqrt(si_sq * sj_sq);
                    dudr_spin = -TR(o, T_DJ, rij0) * (si_dot_sj - si_times_sj);
                    double Jij_half = TR(o, T_J, rij0) / 2e0;
                    double J_times = Jij_half * si_times_sj;
                    double J_dot = -Jij_half * si_dot_sj;
                    o->me0[a] += J_times;
                    o->me[a] += J_dot;
                    o->me0[w] += J_times;
                    o->me[w] += J_dot;
                }
                dudr += dudr_spin;
                double force = -dudr / rij0;
                vec fij = v_times(force, rij);
                o->f[a].x += fij.x; o->f[a].y += fij.y; o->f[a].z += fij.z;
                o->f[w].x -= fij.x; o->f[w].y -= fij.y; o->f[w].z -= fij.z;
                double* st_w = o->stress + 6 * w;
                st_a[0] += fij.x * rij.x; st_a[1] += fij.y * rij.y; st_a[2] += fij.z * rij.z;
                st_a[3] += fij.x * rij.y; st_a[4] += fij.y * rij.z; st_a[5] += fij.z * rij.x;
                st_w[0] += fij.x * rij.x; st_w[1] += fij.y * rij.y; st_w[2] += fij.z * rij.z;
                st_w[3] += fij.x * rij.y; st_w[4] += fij.y * rij.z; st_w[5] += fij.z * rij.x;
                o->vir[a] += -force * rsq;
            }
        }
    }
    double ri = o->rho[a];
    o->pe[a] += TRHO(o, T_F, ri);
    if (o->prm.extfield) o->me[a] -= v_dot(o->s[a], o->hext[a]);  /* calculate_force_energy_CPU.cpp:199-201 */
    o->me[a] += TRHO(o, T_LA, ri) * si_sq + TRHO(o, T_LB, ri) * pow(si_sq, 2) +
                TRHO(o, T_LC, ri) * pow(si_sq, 3) + TRHO(o, T_LD, ri) * pow(si_sq, 4);
}

/* calculate_force_energy_CPU.cpp:220-301 */
void orc_calculate_force_energy(orc_ctx* o) {
    int N = o->natom;
    for (int i = 0; i < N; ++i) {
        o->f[i] = v_zero();
        o->pe[i] = 0e0; o->vir[i] = 0e0;
        for (int k = 0; k < 6; ++k) o->stress[6 * i + k] = 0e0;
        o->me[i] = 0e0; o->me0[i] = 0e0;
        o->sum_rij_m1[i] = 0e0; o->sum_rij_m2[i] = 0e0;
    }
    for (int g = 0; g < o->ngroups; ++g)
        for (int j = 0; j < o->group_size[g]; ++j) {
            int c = o->group_members[g * o->max_members + j];
            for (int a = o->head[c]; a != -1; a = o->next[a]) inner_loop(o, a);
        }
    /* calculate_force_energy_CPU.cpp:266-283 */
    if (o->prm.localvol) {
        double sum_volume = 0e0;
        for (int i = 0; i < N; ++i) {
            double local_radius = 0.5e0 * o->sum_rij_m1[i] / o->sum_rij_m2[i];
            o->local_volume[i] = 4e0 * PI / 3e0 * pow(local_radius, 3e0);
            sum_volume += o->local_volume[i];
        }
        double volume_factor = o->box_volume / sum_volume;
        for (int i = 0; i < N; ++i) o->local_volume[i] *= volume_factor;
    } else {
        for (int i = 0; i < N; ++i) o->local_volume[i] = o->box_volume / N;
    }
    double m = o->prm.atmass;
    for (int i = 0; i < N; ++i) {
        double* st = o->stress + 6 * i;
        vec p = o->p[i];
        double lv = o->local_volume[i];
        st[0] = (pow(p.x, 2) / m + st[0] / 2e0) / lv;
        st[1] = (pow(p.y, 2) / m + st[1] / 2e0) / lv;
        st[2] = (pow(p.z, 2) / m + st[2] / 2e0) / lv;
        st[3] = ((p.x * p.y) / m + st[3] / 2e0) / lv;
        st[4] = ((p.y * p.z) / m + st[4] / 2e0) / lv;
        st[5] = ((p.z * p.x) / m + st[5] / 2e0) / lv;
    }
    double vir = 0e0;
    for (int i = 0; i < N; ++i) vir += o->vir[i];
    o->virial = vir;
}

/* ------------------------------------------------------------------------ */
/* spins                                                                     */

/* calculate_spin_CPU.cpp:201-235: 27-cell scan (neighbours 0..25, then own) */
static void inner_spin(orc_ctx* o, int a) {
    int cc = o->cell_new[a];
    for (int i = 0; i <= 26; ++i) {
        int wc = (i == 26) ? cc : o->neigh[26 * cc + i];
        for (int w = o->head[wc]; w != -1; w = o->next[w]) {
            vec rij = v_sub(o->r[a], o->r[w]);
            find_image(o, &rij);
            double rsq = v_sq(rij);
            if (rsq < o->rcut_mag_sq && a != w) {
                double rij0 = sqrt(rsq);
                double J = TR(o, T_J, rij0);
                o->heff_h[a] = v_add(o->heff_h[a], v_times(J, o->s[w]));
            }
        }
    }
}

/* calculate_spin_CPU.cpp:237-274 */
static void inner_sum_Jij_sj(orc_ctx* o, int a) {
    int cc = o->cell_new[a];
    for (int i = 0; i <= 26; ++i) {
        int wc = (i == 26) ? cc : o->neigh[26 * cc + i];
        for (int w = o->head[wc]; w != -1; w = o->next[w]) {
            vec rij = v_sub(o->r[a], o->r[w]);
            find_image(o, &rij);
            double rsq = v_sq(rij);
            if (rsq < o->rcut_mag_sq && a != w) {
                double rij0 = sqrt(rsq);
                double J = TR(o, T_J, rij0);
                double sj = v_length(o->s[w]);
                o->sum_jij_sj[a] += J * sj;
            }
        }
    }
}

/* core_ds_CPU.cpp:90-127: rotation of s about Heff by |Heff| dt / hbar */
static vec spin_rotation(vec Heff, vec s, double dt) {
    vec omega = v_divide(Heff, -HBAR);
    double omega0 = v_length(omega);
    if (omega0 > 0e0) omega = v_divide(omega, omega0);
    else omega = v_zero();
    double o12 = omega.x * omega.y, o23 = omega.y * omega.z, o13 = omega.x * omega.z;
    double o1s = omega.x * omega.x, o2s = omega.y * omega.y, o3s = omega.z * omega.z;
    double A = sin(omega0 * dt);
    double B = 1e0 - cos(omega0 * dt);
    vec t;
    t.x = s.x + (s.x * B * (-o2s - o3s) + s.y * (B * o12 - A * omega.z) + s.z * (A * omega.y + B * o13));
    t.y = s.y + (s.y * B * (-o1s - o3s) + s.z * (B * o23 - A * omega.x) + s.x * (A * omega.z + B * o12));
    t.z = s.z + (s.z * B * (-o1s - o2s) + s.x * (B * o13 - A * omega.y) + s.y * (A * omega.x + B * o23));
    return t;
}

/* Landau field plus the Heisenberg correction (calculate_spin_CPU.cpp:131-140) */
static inline vec landau_field(orc_ctx* o, int a, vec s, double A, double B, double C, double D) {
    double s_sq = v_sq(s);
    vec L = v_times(-(2e0 * A + 4e0 * B * s_sq + 6e0 * C * pow(s_sq, 2) + 8e0 * D * pow(s_sq, 3)), s);
    double s0 = sqrt(s_sq);
    vec HC = v_zero();
    if (s0 > 0e0) HC = v_times(-o->sum_jij_sj[a] / s0, s);
    o->heff_hc[a] = HC;
    return v_add(L, HC);
}

/* Effective field at the current spins: inner_spin + inner_sum_Jij_sj +
 * landau_field (calculate_spin_CPU.cpp:111-140,201-274) without the spin
 * update; overwrites the heff_h, sum_jij_sj and heff_hc scratch of the atoms. */
void orc_fields(orc_ctx* o, double* h) {
    for (int a = 0; a < o->natom; ++a) {
        o->heff_h[a] = hext_of(o, a);
        inner_spin(o, a);
        o->sum_jij_sj[a] = 0e0;
        inner_sum_Jij_sj(o, a);
        double ri = o->rho[a];
        vec L = landau_field(o, a, o->s[a], TRHO(o, T_LA, ri), TRHO(o, T_LB, ri), TRHO(o, T_LC, ri),
                             TRHO(o, T_LD, ri));
        vec t = v_add(o->heff_h[a], L);
        h[3 * a] = t.x; h[3 * a + 1] = t.y; h[3 * a + 2] = t.z;
    }
}

/* calculate_spin_CPU.cpp:33-199 (SLDHL) */
static void calculate_spin(orc_ctx* o, int a, double dt) {
    o->heff_h[a] = hext_of(o, a);
    o->heff_l[a] = v_zero();
    inner_spin(o, a);
    if (!o->prm.spinlang) {
        o->s[a] = spin_rotation(o->heff_h[a], o->s[a], dt);
    } else {
        double dt_half = dt / 2e0;
        double dt_quad = dt / 4e0;
        double g = o->prm.gamma_S_HL;
        o->s[a] = spin_rotation(o->heff_h[a], o->s[a], dt_half);
        double ri = o->rho[a];
        double A = TRHO(o, T_LA, ri), B = TRHO(o, T_LB, ri), C = TRHO(o, T_LC, ri), D = TRHO(o, T_LD, ri);
        /* RK2 (part 2) */
        o->sum_jij_sj[a] = 0e0;
        inner_sum_Jij_sj(o, a);
        vec H = o->heff_h[a];
        o->heff_l[a] = landau_field(o, a, o->s[a], A, B, C, D);
        vec s_temp = v_add(o->s[a], v_times(g * dt_quad, v_add(H, o->heff_l[a])));
        o->heff_l[a] = landau_field(o, a, s_temp, A, B, C, D);
        o->s[a] = v_add(o->s[a], v_times(g * dt_half, v_add(H, o->heff_l[a])));
        /* noise (part 3) */
        double T = o->prm.eltemp ? o->Te[o->cell_new[a]] : o->prm.temperature;
        double random_S = sqrt(2e0 * T * g / dt);
        double nz[3];
        noise3(o, a, nz);
        vec dS = {random_S * nz[0], random_S * nz[1], random_S * nz[2]};
        o->s[a] = v_add(o->s[a], v_times(dt, dS));
        /* RK2 (part 4) */
        o->heff_l[a] = landau_field(o, a, o->s[a], A, B, C, D);
        s_temp = v_add(o->s[a], v_times(g * dt_quad, v_add(H, o->heff_l[a])));
        o->heff_l[a] = landau_field(o, a, s_temp, A, B, C, D);
        o->s[a] = v_add(o->s[a], v_times(g * dt_half, v_add(H, o->heff_l[a])));
        /* part 5 */
        o->s[a] = spin_rotation(H, o->s[a], dt_half);
    }
    o->s0[a] = v_length(o->s[a]);
}

/* core_ds_CPU.cpp:35-83: forward sweep over groups (head to tail), then
 * backward (tail to head). Each sweep is one noise pass. */
void orc_core_ds(orc_ctx* o, double dt) {
    double dt_half = dt / 2e0;
    if (o->prm.magmom) /* core_ds_CPU.cpp:37-45 */
        for (int i = 0; i < o->natom; ++i) {
            o->s[i] = v_divide(o->m[i], -EL_G);
            o->s0[i] = v_length(o->s[i]);
        }
    for (int g = 0; g < o->ngroups; ++g) {
        for (int j = 0; j < o->group_size[g]; ++j) {
            int c = o->group_members[g * o->max_members + j];
            for (int a = o->head[c]; a != -1; a = o->next[a]) calculate_spin(o, a, dt_half);
        }
    }
    o->noise_call++;
    for (int g = o->ngroups - 1; g >= 0; --g) {
        for (int j = 0; j < o->group_size[g]; ++j) {
            int c = o->group_members[g * o->max_members + j];
            for (int a = o->tail[c]; a != -1; a = o->prev[a]) calculate_spin(o, a, dt_half);
        }
    }
    o->noise_call++;
    if (o->prm.magmom) /* core_ds_CPU.cpp:73-81 */
        for (int i = 0; i < o->natom; ++i) o->m[i] = v_times(-EL_G, o->s[i]);
}

/* ------------------------------------------------------------------------ */
/* lattice                                                                   */

/* core_dp_CPU.cpp:268-312 */
static void rescale_momentum(orc_ctx* o) {
    int N = o->natom;
    double m = o->prm.atmass;
    double ax = 0e0, ay = 0e0, az = 0e0;
    for (int i = 0; i < N; ++i) { ax += o->p[i].x; ay += o->p[i].y; az += o->p[i].z; }
    ax /= N; ay /= N; az /= N;
    double ke_old = 0e0, ke_new = 0e0;
    for (int i = 0; i < N; ++i) {
        o->ke[i] = v_sq(o->p[i]) / 2e0 / m;
        ke_old += o->ke[i];
        o->p[i].x -= ax; o->p[i].y -= ay; o->p[i].z -= az;
        o->ke[i] = v_sq(o->p[i]) / 2e0 / m;
        ke_new += o->ke[i];
    }
    double factor = sqrt(ke_old / ke_new);
    if (ke_new < 1e-10) factor = 1e0;
    for (int i = 0; i < N; ++i) o->p[i] = v_times(factor, o->p[i]);
}

/* core_dp_CPU.cpp:90-170 */
static void core_dp_A(orc_ctx* o, double dt) {
    double gl = o->gamma_L;
    if (o->prm.localcolmot) {
        for (int g = 0; g < o->ngroups; ++g)
            for (int j = 0; j < o->group_size[g]; ++j) {
                int c = o->group_members[g * o->max_members + j];
                int a = o->head[c];
                int active = 0;
                if (a != -1) { o->ave_fluct[c] = v_zero(); active = 1; }
                for (; a != -1; a = o->next[a]) {
                    double T = o->prm.eltemp ? o->Te[c] : o->prm.temperature;
                    vec ff = fluct_force(o, a, T, gl, dt);
                    o->p[a] = v_add(o->p[a], v_times(dt, v_add(dp_force(o, a), ff)));
                    o->ave_fluct[c] = v_add(o->ave_fluct[c], ff);
                }
                if (active) o->ave_fluct[c] = v_divide(o->ave_fluct[c], o->count[c]);
            }
        for (int i = 0; i < o->natom; ++i)
            o->p[i] = v_sub(o->p[i], v_times(dt, o->ave_fluct[o->cell_new[i]]));
    } else {
        for (int i = 0; i < o->natom; ++i) {
            double T = o->prm.eltemp ? o->Te[o->cell_new[i]] : o->prm.temperature;
            vec ff = fluct_force(o, i, T, gl, dt);
            o->p[i] = v_add(o->p[i], v_times(dt, v_add(dp_force(o, i), ff)));
        }
    }
    o->noise_call++;
}

/* core_dp_CPU.cpp:172-192 */
static void core_dp_B(orc_ctx* o, double dt) {
    double gl = o->gamma_L, m = o->prm.atmass;
    for (int i = 0; i < o->natom; ++i) {
        double e;
        if (o->prm.localcolmot) e = exp(-gl / m * dt * (1e0 - 1e0 / o->count[o->cell_new[i]]));
        else e = exp(-gl / m * dt);
        o->p[i] = v_times(e, o->p[i]);
    }
}

/* core_dp_CPU.cpp:211-264: per cell, p_i += gamma/m dt/n (sum_p - p_i),
 * head to tail (C1, groups forward) or tail to head (C2, groups backward). */
static void core_dp_C(orc_ctx* o, double dt, int backward) {
    double gl = o->gamma_L, m = o->prm.atmass;
    for (int gg = 0; gg < o->ngroups; ++gg) {
        int g = backward ? o->ngroups - 1 - gg : gg;
        for (int j = 0; j < o->group_size[g]; ++j) {
            int c = o->group_members[g * o->max_members + j];
            int a = backward ? o->tail[c] : o->head[c];
            while (a != -1) {
                vec sum_p = v_zero();
                for (int w = o->head[c]; w != -1; w = o->next[w]) sum_p = v_add(sum_p, o->p[w]);
                double factor = gl / m * dt / o->count[o->cell_new[a]];
                o->p[a] = v_add(o->p[a], v_times(factor, v_sub(sum_p, o->p[a])));
                a = backward ? o->prev[a] : o->next[a];
            }
        }
    }
}

/* core_dp_CPU.cpp:51-85 */
void orc_core_dp(orc_ctx* o, double dt) {
    if (o->prm.lattlang) {
        if (o->prm.localcolmot) core_dp_C(o, dt / 2e0, 0);
        core_dp_B(o, dt / 2e0);
        core_dp_A(o, dt);
        core_dp_B(o, dt / 2e0);
        if (o->prm.localcolmot) core_dp_C(o, dt / 2e0, 1);
    } else {
        for (int i = 0; i < o->natom; ++i) o->p[i] = v_add(o->p[i], v_times(dt, dp_force(o, i)));
    }
    rescale_momentum(o);
    for (int i = 0; i < o->natom; ++i) o->ke[i] = v_sq(o->p[i]) / 2e0 / o->prm.atmass;
}

/* core_dr_CPU.cpp:33-44 */
void orc_core_dr(orc_ctx* o, double dt) {
    double k = dt / o->prm.atmass;
    for (int i = 0; i < o->natom; ++i) {
        o->r[i] = v_add(o->r[i], v_times(k, o->p[i]));
        periodic(o, &o->r[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* electron subsystem                                                        */

/* calculate_temperature_CPU.cpp:35-148 */
void orc_calculate_temperature(orc_ctx* o) {
    double m = o->prm.atmass, gl = o->gamma_L;
    double apc = (double)o->natom / (double)o->ncells;
    for (int c = 0; c < o->ncells; ++c) {
        double sum_ke = 0e0;
        if (o->prm.localcolmot) o->ave_p[c] = v_zero();
        o->sum_R_up[c] = 0e0; o->sum_R_dn[c] = 0e0;
        o->sum_L_up[c] = 0e0; o->sum_L_dn[c] = 0e0;
        for (int a = o->head[c]; a != -1; a = o->next[a]) {
            sum_ke += o->ke[a];
            if (o->prm.localcolmot) o->ave_p[c] = v_add(o->ave_p[c], o->p[a]);
            o->heff_h[a] = hext_of(o, a);
            inner_spin(o, a);
            o->sum_R_up[c] += v_sq(v_cross(o->s[a], o->heff_h[a]));
            o->sum_R_dn[c] += v_dot(o->s[a], o->heff_h[a]);
            o->heff_l[a] = v_zero();
            double ri = o->rho[a];
            double A = TRHO(o, T_LA, ri), B = TRHO(o, T_LB, ri), C = TRHO(o, T_LC, ri), D = TRHO(o, T_LD, ri);
            double s_sq = v_sq(o->s[a]);
            o->s0[a] = v_length(o->s[a]);
            o->heff_l[a] = v_times(-(2e0 * A + 4e0 * B * s_sq + 6e0 * C * pow(s_sq, 2) + 8e0 * D * pow(s_sq, 3)), o->s[a]);
            o->sum_jij_sj[a] = 0e0;
            inner_sum_Jij_sj(o, a);
            o->heff_hc[a] = v_times(-o->sum_jij_sj[a] / o->s0[a], o->s[a]);
            o->heff_l[a] = v_add(o->heff_l[a], o->heff_hc[a]);
            o->sum_L_up[c] += v_sq(v_add(o->heff_h[a], o->heff_l[a]));
            o->sum_L_dn[c] += 6e0 * A + 20e0 * B * s_sq + 42e0 * C * pow(s_sq, 2) + 72e0 * D * pow(s_sq, 3);
            o->sum_L_dn[c] += 2e0 * o->sum_jij_sj[a] / o->s0[a];
        }
        double n = (double)o->count[c];
        if (o->prm.localcolmot) {
            o->ave_p[c] = v_divide(o->ave_p[c], o->count[c]);
            sum_ke -= o->count[c] * v_sq(o->ave_p[c]) / 2e0 / m;
            o->Tl[c] = 2e0 / 3e0 * sum_ke / (n - 1e0) / BOLTZ;
            o->Gel[c] = 3e0 * (n - 1e0) * BOLTZ * gl / m / apc;
        } else {
            o->Tl[c] = sum_ke * 2e0 / 3e0 / o->count[c] / BOLTZ;
            o->Gel[c] = 3e0 * o->count[c] * BOLTZ * gl / m / apc;
        }
        o->Ts_R[c] = o->sum_R_up[c] / o->sum_R_dn[c] / 2e0 / BOLTZ;
        o->Ts_L[c] = o->sum_L_up[c] / o->sum_L_dn[c] / BOLTZ;
        o->Ges[c] = BOLTZ * o->prm.gamma_S_HL * o->sum_L_dn[c] / apc;
    }
}

/* core_dTe_CPU.cpp:66-93 */
static void core_dTe_A(orc_ctx* o, double dt) {
    for (int c = 0; c < o->ncells; ++c) {
        o->Ee[c] = Te_to_Ee(o->Te[c]);
        double dE = 0e0;
        if (o->prm.lattlang) dE += o->Gel[c] * o->Tl[c];
        if (o->prm.spinlang) dE += o->Ges[c] * o->Ts_L[c];
        o->Ee[c] += dt * dE;
        o->Te[c] = Ee_to_Te(o->Ee[c]);
    }
}

/* core_dTe_CPU.cpp:95-129 (RK4) */
static void core_dTe_B(orc_ctx* o, double dt) {
    for (int c = 0; c < o->ncells; ++c) {
        double G = 0e0;
        if (o->prm.lattlang) G += o->Gel[c];
        if (o->prm.spinlang) G += o->Ges[c];
        G *= dt / BOLTZ;
        double T0 = o->Te[c];
        double k1 = T0 * (exp(-G / Ce(T0)) - 1e0) / dt;
        double T1 = T0 + k1 * dt / 2e0;
        double k2 = T1 * (exp(-G / Ce(T1)) - 1e0) / dt;
        double T2 = T0 + k2 * dt / 2e0;
        double k3 = T2 * (exp(-G / Ce(T2)) - 1e0) / dt;
        double T3 = T0 + k3 * dt;
        double k4 = T3 * (exp(-G / Ce(T3)) - 1e0) / dt;
        o->Te[c] += dt / 6e0 * (k1 + 2e0 * k2 + 2e0 * k3 + k4);
    }
}

static inline void dTe_stencil(orc_ctx* o, int c, double pref, double dx2, double dy2, double dz2) {
    const int* n = o->neigh + 26 * c;
    double* Te = o->Te;
    o->Ee[c] = Te_to_Ee(Te[c]);
    o->Ee[c] += pref * ((Te[n[0]] + Te[n[13]] - 2e0 * Te[c]) / dx2 +
                        (Te[n[2]] + Te[n[14]] - 2e0 * Te[c]) / dy2 +
                        (Te[n[12]] + Te[n[15]] - 2e0 * Te[c]) / dz2);
    Te[c] = Ee_to_Te(o->Ee[c]);
}

/* core_dTe_CPU.cpp:131-217 */
static void core_dTe_C(orc_ctx* o, double dt) {
    double start = 0e0;
    for (int c = 0; c < o->ncells; ++c) { o->Ee[c] = Te_to_Ee(o->Te[c]); start += o->Ee[c]; }
    double kappa = o->prm.kappa_e / EVTOJ / 1e10;
    double dx2 = pow(o->box_length.x / o->nx, 2);
    double dy2 = pow(o->box_length.y / o->ny, 2);
    double dz2 = pow(o->box_length.z / o->nz, 2);
    double atomic_volume = o->box_volume / (double)o->natom;
    double volume_factor = atomic_volume / ((double)o->natom / (double)o->ncells);
    double pref = dt / 2e0 * kappa * volume_factor / BOLTZ;
    for (int g = 0; g < o->ngroups; ++g)
        for (int j = 0; j < o->group_size[g]; ++j)
            dTe_stencil(o, o->group_members[g * o->max_members + j], pref, dx2, dy2, dz2);
    for (int g = o->ngroups - 1; g >= 0; --g)
        for (int j = 0; j < o->group_size[g]; ++j)
            dTe_stencil(o, o->group_members[g * o->max_members + j], pref, dx2, dy2, dz2);
    double end = 0e0;
    for (int c = 0; c < o->ncells; ++c) { o->Ee[c] = Te_to_Ee(o->Te[c]); end += o->Ee[c]; }
    double f = start / end;
    for (int c = 0; c < o->ncells; ++c) { o->Ee[c] *= f; o->Te[c] = Ee_to_Te(o->Ee[c]); }
}

/* core_dTe_CPU.cpp:47-60, nloop = 50 */
void orc_core_dTe(orc_ctx* o, double dt) {
    int nloop = 50;
    double dtl = dt / nloop;
    for (int i = 0; i < nloop; ++i) {
        core_dTe_A(o, dtl / 2e0);
        core_dTe_B(o, dtl / 2e0);
        core_dTe_C(o, dtl);
        core_dTe_B(o, dtl / 2e0);
        core_dTe_A(o, dtl / 2e0);
    }
}

/* ------------------------------------------------------------------------ */
/* observables                                                               */

/* check.cpp:31-47 with check_energy_CPU.cpp:33-111,
 * check_temperature_CPU.cpp:33-134 and check_spin_CPU.cpp:33-56 */
void orc_check(orc_ctx* o, int step, sld_obs* out) {
    int N = o->natom;
    sld_obs ob;
    memset(&ob, 0, sizeof(ob));
    ob.step = step;
    ob.total_time = o->total_time;
    double pe = 0e0, ke = 0e0, me = 0e0, me0 = 0e0;
    for (int i = 0; i < N; ++i) { pe += o->pe[i]; ke += o->ke[i]; }
    pe /= N; ke /= N;
    for (int i = 0; i < N; ++i) { me += o->me[i]; me0 += o->me0[i]; }
    me /= N; me0 /= N;
    ob.ave_pe = pe; ob.ave_ke = ke; ob.ave_me = me; ob.ave_me0 = me0;
    if (o->prm.eltemp) {
        double ee = 0e0;
        double apc = (double)N / (double)o->ncells;
        for (int c = 0; c < o->ncells; ++c) { o->Ee[c] = Te_to_Ee(o->Te[c]); ee += apc * o->Ee[c]; }
        ee /= N;
        ob.ave_ee = ee;
        if (o->prm.renormalize_energy) {
            double err = 0e0;
            if (step == -1) {
                double e0 = 0e0;
                e0 += pe; e0 += ke; e0 += me; e0 += me0; e0 += ee;
                o->initial_ave_energy = e0;
            } else {
                double cur = 0e0;
                cur += pe; cur += ke; cur += me; cur += me0; cur += ee;
                err = cur - o->initial_ave_energy;
            }
            for (int c = 0; c < o->ncells; ++c) { o->Ee[c] -= err; o->Te[c] = Ee_to_Te(o->Ee[c]); }
        }
    }
    double sk = 0e0;
    for (int i = 0; i < N; ++i) sk += o->ke[i];
    ob.tl = sk * 2e0 / 3e0 / N / BOLTZ;
    double rup = 0e0, rdn = 0e0, lup = 0e0, ldn = 0e0;
    if (o->prm.eltemp) {
        for (int c = 0; c < o->ncells; ++c) { rup += o->sum_R_up[c]; rdn += o->sum_R_dn[c]; }
        for (int c = 0; c < o->ncells; ++c) { lup += o->sum_L_up[c]; ldn += o->sum_L_dn[c]; }
    } else {
        for (int i = 0; i < N; ++i) {
            o->heff_h[i] = hext_of(o, i);
            inner_spin(o, i);
            rup += v_sq(v_cross(o->s[i], o->heff_h[i]));
            rdn += v_dot(o->s[i], o->heff_h[i]);
        }
        for (int i = 0; i < N; ++i) {
            double ri = o->rho[i];
            double A = TRHO(o, T_LA, ri), B = TRHO(o, T_LB, ri), C = TRHO(o, T_LC, ri), D = TRHO(o, T_LD, ri);
            o->sum_jij_sj[i] = 0e0;
            inner_sum_Jij_sj(o, i);
            ldn += 2e0 * o->sum_jij_sj[i] / v_length(o->s[i]);
            double s_sq = v_sq(o->s[i]);
            o->heff_l[i] = v_zero();
            o->heff_l[i] = v_times(-(2e0 * A + 4e0 * B * s_sq + 6e0 * C * pow(s_sq, 2) + 8e0 * D * pow(s_sq, 3)), o->s[i]);
            o->heff_hc[i] = v_times(-o->sum_jij_sj[i] / v_length(o->s[i]), o->s[i]);
            o->heff_l[i] = v_add(o->heff_l[i], o->heff_hc[i]);
            lup += v_sq(v_add(o->heff_h[i], o->heff_l[i]));
            ldn += 6e0 * A + 20e0 * B * s_sq + 42e0 * C * pow(s_sq, 2) + 72e0 * D * pow(s_sq, 3);
        }
    }
    ob.ts_r = rup / rdn / 2e0 / BOLTZ;
    ob.ts_l = lup / ldn / BOLTZ;
    if (o->prm.eltemp) {
        double te = 0e0;
        for (int c = 0; c < o->ncells; ++c) te += o->Te[c];
        te /= o->ncells;
        te /= BOLTZ;
        ob.te = te;
    }
    double sx = 0e0, sy = 0e0, sz = 0e0;
    for (int i = 0; i < N; ++i) { sx += o->s[i].x; sy += o->s[i].y; sz += o->s[i].z; }
    vec as = {sx, sy, sz};
    as = v_divide(as, N);
    ob.ave_s[0] = as.x; ob.ave_s[1] = as.y; ob.ave_s[2] = as.z;
    ob.ave_s0 = v_length(as);
    vec am = v_times(-EL_G, as);
    ob.ave_m[0] = am.x; ob.ave_m[1] = am.y; ob.ave_m[2] = am.z;
    ob.ave_m0 = v_length(am);
    ob.virial = o->virial;
    /* check_stress_CPU.cpp:35-58, check_pressure_CPU.cpp:35-41 */
    double st[6] = {0e0, 0e0, 0e0, 0e0, 0e0, 0e0};
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < 6; ++k) st[k] += o->stress[6 * i + k];
    for (int k = 0; k < 6; ++k) ob.ave_stress[k] = st[k] * (160.217653e0 / N);
    double density = (double)N / o->box_volume;
    double tmp = 2e0 / 3e0 * sk / N;
    ob.pressure = density * (tmp - o->virial / (3e0 * N));
    ob.pressure *= 160.217653e0;
    memcpy(o->last_stress, ob.ave_stress, sizeof(o->last_stress));
    o->last_pressure = ob.pressure;
    if (out) *out = ob;
}

/* core.cpp:31-76 (SLDHL) */
void orc_core(orc_ctx* o, int step) {
    double dt = o->prm.step;
    orc_links(o);
    if (o->prm.eltemp) orc_core_dTe(o, dt / 2e0);
    orc_core_dp(o, dt / 2e0);
    orc_core_ds(o, dt / 2e0);
    orc_core_dr(o, dt);
    orc_calculate_rho(o);
    orc_core_ds(o, dt / 2e0);
    orc_calculate_force_energy(o);
    orc_core_dp(o, dt / 2e0);
    if (o->prm.eltemp) {
        orc_calculate_temperature(o);
        orc_core_dTe(o, dt / 2e0);
    }
    (void)step;
}

/* spilady.cpp:43-57 with runstep: core, total_time += step, check */
/* scale_step_CPU.cpp:33-97 (single thread): the largest |p|^2 and
 * |Heff_H|^2 of the step just taken set the displacement and rotation
 * estimates; dt shrinks by 0.8 if either exceeds its limit, else grows by 1.05 */
void orc_scale_step(orc_ctx* o) {
    double displace_max = 0e0, phi_max = 0e0;
    for (int i = 0; i < o->natom; ++i) {
        double p_sq = v_sq(o->p[i]);
        if (p_sq > displace_max) displace_max = p_sq;
    }
    displace_max = sqrt(displace_max) / o->prm.atmass * o->prm.step;
    for (int i = 0; i < o->natom; ++i) {
        double omega_sq = v_sq(o->heff_h[i]);
        if (omega_sq > phi_max) phi_max = omega_sq;
    }
    phi_max = sqrt(phi_max) / HBAR * o->prm.step;
    int switch_lattice = displace_max > o->prm.displace_limit ? 1 : 0;
    int switch_spin = phi_max > o->prm.phi_limit ? 1 : 0;
    if (switch_spin + switch_lattice > 0) o->prm.step *= 0.80;
    else o->prm.step *= 1.05;
}

double orc_get_step(const orc_ctx* o) { return o->prm.step; }

/* box_length, volume and inverse of a new box (scale_pressure_CPU.cpp:47-55) */
static void new_box(orc_ctx* o, const double* d) {
    memcpy(o->d, d, sizeof(o->d));
    inverse_box(o->d, o->inv);
    o->box_length.x = fabs(o->d[0]);
    o->box_length.y = sqrt(o->d[1] * o->d[1] + o->d[2] * o->d[2]);
    o->box_length.z = sqrt(o->d[3] * o->d[3] + o->d[4] * o->d[4] + o->d[5] * o->d[5]);
    o->box_volume = o->d[0] * o->d[2] * o->d[5];
}

/* scale_stress_CPU.cpp:33-86: each box entry moves with the difference
 * between the last checked average stress and the target; the atoms keep
 * their coordinates in the old box's frame. ave_stress is 11 22 33 12 23 31,
 * the targets are xx yx yy zx zy zz. */
static void scale_stress(orc_ctx* o) {
    double delta_time = o->total_time - o->last_t_stress;
    o->last_t_stress = o->total_time;
    double pre = delta_time / o->baro.baro_damping_time / 100e0;
    const double *a = o->last_stress, *t = o->baro.stress;
    double fxx = pre * (a[0] - t[0]), fyx = pre * (a[3] - t[1]), fyy = pre * (a[1] - t[2]);
    double fzx = pre * (a[5] - t[3]), fzy = pre * (a[4] - t[4]), fzz = pre * (a[2] - t[5]);
    double d[6], I[6];
    memcpy(d, o->d, sizeof(d));
    memcpy(I, o->inv, sizeof(I));
    d[0] = (1e0 + fxx) * d[0];
    d[1] = (1e0 + fxx) * d[1] + fyx * (d[2] + d[0]) / 2e0;
    d[2] = (1e0 + fyy) * d[2];
    d[3] = (1e0 + fxx) * d[3] + fzx * (d[5] + d[0]) / 2e0;
    d[4] = (1e0 + fyy) * d[4] + fzy * (d[5] + d[2]) / 2e0;
    d[5] = (1e0 + fzz) * d[5];
    for (int i = 0; i < o->natom; ++i) {
        vec r = o->r[i], q;
        q.x = I[0] * r.x + I[1] * r.y + I[3] * r.z;
        q.y = I[2] * r.y + I[4] * r.z;
        q.z = I[5] * r.z;
        o->r[i].x = d[0] * q.x + d[1] * q.y + d[3] * q.z;
        o->r[i].y = d[2] * q.y + d[4] * q.z;
        o->r[i].z = d[5] * q.z;
    }
    new_box(o, d);
}

/* scale_pressure_CPU.cpp:33-61: isotropic factor from the last checked pressure */
static void scale_pressure(orc_ctx* o) {
    double delta_time = o->total_time - o->last_t_pressure;
    o->last_t_pressure = o->total_time;
    double factor = pow(1e0 + delta_time / o->baro.baro_damping_time * (o->last_pressure - o->baro.pressure) / 100e0,
                        1e0 / 3e0);
    double d[6];
    for (int k = 0; k < 6; ++k) d[k] = o->d[k] * factor;
    new_box(o, d);
    for (int i = 0; i < o->natom; ++i) o->r[i] = v_times(factor, o->r[i]);
}

void orc_set_barostat(orc_ctx* o, const sld_barostat* b) { o->baro = *b; }

/* scale.cpp:31-44 */
void orc_scale(orc_ctx* o, int step) {
    if (o->baro.stress_on && (step + 1) % o->baro.interval_of_scale_stress == 0) scale_stress(o);
    if (o->baro.pressure_on && (step + 1) % o->baro.interval_of_scale_pressure == 0) scale_pressure(o);
    if (o->prm.changestep) orc_scale_step(o);
}

void orc_get_box(const orc_ctx* o, double* box) { memcpy(box, o->d, sizeof(o->d)); }

/* the quantum thermostat's filter (2 nfrequency_quantum taps); 0 without it */
int orc_quantum_H(const orc_ctx* o, double* out) {
    if (!o->qH) return 0;
    if (out) memcpy(out, o->qH, sizeof(double) * o->qn_nf2);
    return o->qn_nf2;
}

/* spilady.cpp:43-57: core, time, check on printed steps, scale */
int orc_step(orc_ctx* o, int first_step, int nsteps, sld_obs* obs) {
    int k = 0;
    int every = o->prm.interval_of_print_out > 0 ? o->prm.interval_of_print_out : 1;
    for (int s = first_step; s < first_step + nsteps; ++s) {
        orc_core(o, s);
        o->total_time += o->prm.step;
        if ((s + 1) % every == 0) {
            orc_check(o, s, obs ? obs + k : NULL);
            ++k;
        }
        orc_scale(o, s);
    }
    return k;
}

/* ------------------------------------------------------------------------ */
/* set-up and state access                                                   */

#define ALLOCV(n) ((vec*)calloc((size_t)(n), sizeof(vec)))
#define ALLOCD(n) ((double*)calloc((size_t)(n), sizeof(double)))
#define ALLOCI(n) ((int*)calloc((size_t)(n), sizeof(int)))

static void load3(vec* dst, const double* src, int n) {
    if (!src) return;
    for (int i = 0; i < n; ++i) { dst[i].x = src[3 * i]; dst[i].y = src[3 * i + 1]; dst[i].z = src[3 * i + 2]; }
}
static void store3(double* dst, const vec* src, int n) {
    if (!dst) return;
    for (int i = 0; i < n; ++i) { dst[3 * i] = src[i].x; dst[3 * i + 1] = src[i].y; dst[3 * i + 2] = src[i].z; }
}
static void load1(double* dst, const double* src, int n) { if (src) memcpy(dst, src, sizeof(double) * n); }
static void store1(double* dst, const double* src, int n) { if (dst) memcpy(dst, src, sizeof(double) * n); }

/* initial.cpp:31-106 with readconf (read_config.cpp:110-145 for the box) */
/* magmom: the caller's spin field holds the moments m; the spins follow as
 * s = m / (-g) (read_config.cpp:136-139) */
static void magmom_from_input(orc_ctx* o) {
    if (!o->prm.magmom) return;
    for (int i = 0; i < o->natom; ++i) {
        o->m[i] = o->s[i];
        o->s[i] = v_divide(o->m[i], -EL_G);
    }
}

orc_ctx* orc_create(const sld_params* p, const sld_atoms* in, int compute_initial_forces, int rng_mode) {
    orc_ctx* o = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    o->prm = *p;
    if (o->prm.rcut_max > o->prm.min_length_link_cell) o->prm.min_length_link_cell = o->prm.rcut_max + 0.001;
    o->rng_mode = rng_mode;
    int N = o->natom = p->natom;
    memcpy(o->d, p->box, sizeof(o->d));
    inverse_box(o->d, o->inv);
    o->box_length.x = fabs(o->d[0]);
    o->box_length.y = sqrt(o->d[1] * o->d[1] + o->d[2] * o->d[2]);
    o->box_length.z = sqrt(o->d[3] * o->d[3] + o->d[4] * o->d[4] + o->d[5] * o->d[5]);
    o->box_volume = o->d[0] * o->d[2] * o->d[5];
    o->gamma_L = p->gamma_L_over_mass * p->atmass;
    o->rcut_pot_sq = pow(p->rcut_pot, 2);
    o->rcut_mag_sq = pow(p->rcut_mag, 2);
    o->rcut_max_sq = pow(p->rcut_max, 2);
    potential_table(o);

    o->r = ALLOCV(N); o->p = ALLOCV(N); o->f = ALLOCV(N); o->s = ALLOCV(N);
    o->heff_h = ALLOCV(N); o->heff_l = ALLOCV(N); o->heff_hc = ALLOCV(N);
    o->hext = ALLOCV(N); o->fext = ALLOCV(N);
    o->s0 = ALLOCD(N); o->rho = ALLOCD(N); o->pe = ALLOCD(N); o->ke = ALLOCD(N);
    o->me = ALLOCD(N); o->me0 = ALLOCD(N); o->vir = ALLOCD(N); o->stress = ALLOCD(6 * N);
    o->sum_jij_sj = ALLOCD(N); o->local_volume = ALLOCD(N);
    o->sum_rij_m1 = ALLOCD(N); o->sum_rij_m2 = ALLOCD(N);
    o->next = ALLOCI(N); o->prev = ALLOCI(N); o->cell_old = ALLOCI(N); o->cell_new = ALLOCI(N);
    load3(o->r, in->r, N); load3(o->p, in->p, N); load3(o->f, in->f, N); load3(o->s, in->s, N);
    load3(o->heff_h, in->heff_h, N); load3(o->heff_l, in->heff_l, N);
    load1(o->rho, in->rho, N); load1(o->pe, in->pe, N); load1(o->ke, in->ke, N);
    load1(o->me, in->me, N); load1(o->me0, in->me0, N);
    if (in->stress) memcpy(o->stress, in->stress, sizeof(double) * 6 * N);
    o->m = ALLOCV(N);
    magmom_from_input(o);
    for (int i = 0; i < N; ++i) o->s0[i] = v_length(o->s[i]);

    map_cells(o);
    int C = o->ncells;
    o->head = ALLOCI(C); o->tail = ALLOCI(C); o->count = ALLOCI(C); o->type = ALLOCI(C);
    o->Te = ALLOCD(C); o->Ee = ALLOCD(C); o->Tl = ALLOCD(C); o->Gel = ALLOCD(C); o->Ges = ALLOCD(C);
    o->Ts_R = ALLOCD(C); o->Ts_L = ALLOCD(C);
    o->sum_R_up = ALLOCD(C); o->sum_R_dn = ALLOCD(C); o->sum_L_up = ALLOCD(C); o->sum_L_dn = ALLOCD(C);
    o->ave_p = ALLOCV(C); o->ave_fluct = ALLOCV(C);
    if (o->prm.eltemp)   /* map_cells.cpp:128-159 */
        for (int c = 0; c < C; ++c) { o->Te[c] = p->temperature; o->Ee[c] = Te_to_Ee(o->Te[c]); }
    initial_links(o);
    allocate_cells(o);
    if (compute_initial_forces) {
        orc_calculate_rho(o);
        orc_calculate_force_energy(o);
    }
    if (o->prm.eltemp) orc_calculate_temperature(o);
    o->total_time = p->start_time;
    o->last_t_stress = o->last_t_pressure = o->total_time;  /* initial.cpp:93-98 */
    /* random_number_CPU.cpp:46-48 for thread 0 */
    o->m_w = (uint32_t)(0 + (int)p->random_seed);
    o->m_z = (uint32_t)(0 * 3 + (int)p->random_seed * 2);
    o->spare = 0;
    o->noise_call = 0;
    if (o->prm.quantumnoise) quantum_init(o);  /* initial.cpp:102-104 */
    orc_check(o, -1, NULL);
    return o;
}

void orc_destroy(orc_ctx* o) {
    if (!o) return;
    for (int t = 0; t < NTAB; ++t) free(o->tab[t]);
    void* ptrs[] = {o->r, o->p, o->f, o->s, o->heff_h, o->heff_l, o->heff_hc, o->s0, o->rho, o->pe,
                    o->ke, o->me, o->me0, o->vir, o->stress, o->sum_jij_sj, o->local_volume,
                    o->sum_rij_m1, o->sum_rij_m2, o->m, o->next,
                    o->prev, o->cell_old, o->cell_new, o->head, o->tail, o->count, o->neigh, o->type,
                    o->group_size, o->group_members, o->Te, o->Ee, o->Tl, o->Gel, o->Ges, o->Ts_R,
                    o->Ts_L, o->sum_R_up, o->sum_R_dn, o->sum_L_up, o->sum_L_dn, o->ave_p, o->ave_fluct,
                    o->hext, o->fext, o->qH, o->qmem, o->qval, o->qcount};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
    free(o);
}

/* external field (eV) and force (eV/A) per atom, [natom][3]; NULL: unchanged */
void orc_set_external(orc_ctx* o, const double* hext, const double* fext) {
    for (int i = 0; i < o->natom; ++i) {
        if (hext) { o->hext[i].x = hext[3 * i]; o->hext[i].y = hext[3 * i + 1]; o->hext[i].z = hext[3 * i + 2]; }
        if (fext) { o->fext[i].x = fext[3 * i]; o->fext[i].y = fext[3 * i + 1]; o->fext[i].z = fext[3 * i + 2]; }
    }
}

void orc_download(const orc_ctx* o, sld_atoms* out) {
    int N = o->natom;
    store3(out->r, o->r, N); store3(out->p, o->p, N); store3(out->f, o->f, N);
    store3(out->s, o->prm.magmom ? o->m : o->s, N);
    store3(out->heff_h, o->heff_h, N); store3(out->heff_l, o->heff_l, N);
    store1(out->rho, o->rho, N); store1(out->pe, o->pe, N); store1(out->ke, o->ke, N);
    store1(out->me, o->me, N); store1(out->me0, o->me0, N);
    if (out->stress) memcpy(out->stress, o->stress, sizeof(double) * 6 * N);
    store1(out->vir, o->vir, N); store1(out->sum_jij_sj, o->sum_jij_sj, N);
    store1(out->local_volume, o->local_volume, N);
}

void orc_upload(orc_ctx* o, const sld_atoms* in) {
    int N = o->natom;
    load3(o->r, in->r, N); load3(o->p, in->p, N); load3(o->s, in->s, N);
    if (in->s) magmom_from_input(o);
    load3(o->f, in->f, N); load1(o->rho, in->rho, N);
    for (int i = 0; i < N; ++i) o->s0[i] = v_length(o->s[i]);
}

void orc_get_cells(const orc_ctx* o, sld_cells* out) {
    int C = o->ncells;
    store1(out->te, o->Te, C); store1(out->ee, o->Ee, C); store1(out->tl, o->Tl, C);
    store1(out->ts_r, o->Ts_R, C); store1(out->ts_l, o->Ts_L, C); store1(out->gel, o->Gel, C);
    store1(out->ges, o->Ges, C); store1(out->sum_r_up, o->sum_R_up, C);
    store1(out->sum_r_dn, o->sum_R_dn, C); store1(out->sum_l_up, o->sum_L_up, C);
    store1(out->sum_l_dn, o->sum_L_dn, C);
}

/* Te of every cell and Ee = E(Te), as the readTe start sets them (map_cells.cpp:136-157) */
void orc_set_cells_te(orc_ctx* o, const double* te) {
    memcpy(o->Te, te, sizeof(double) * o->ncells);
    for (int c = 0; c < o->ncells; ++c) o->Ee[c] = Te_to_Ee(o->Te[c]);
}

void orc_get_geometry(const orc_ctx* o, int* ncell3, int* ngroups, int* group_sizes) {
    if (ncell3) { ncell3[0] = o->nx; ncell3[1] = o->ny; ncell3[2] = o->nz; }
    if (ngroups) *ngroups = o->ngroups;
    if (group_sizes) memcpy(group_sizes, o->group_size, sizeof(int) * o->ngroups);
}

int orc_get_groups(const orc_ctx* o, int* members) {
    int k = 0;
    for (int g = 0; g < o->ngroups; ++g)
        for (int j = 0; j < o->group_size[g]; ++j) members[k++] = o->group_members[g * o->max_members + j];
    return k;
}

int orc_get_cell_lists(const orc_ctx* o, int cap, int* counts, int* lists) {
    int mx = 0;
    for (int c = 0; c < o->ncells; ++c) {
        int k = 0;
        for (int a = o->head[c]; a != -1; a = o->next[a]) {
            if (k < cap) lists[c * cap + k] = a;
            ++k;
        }
        counts[c] = k;
        if (k > mx) mx = k;
    }
    return mx;
}

double orc_total_time(const orc_ctx* o) { return o->total_time; }
