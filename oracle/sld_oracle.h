/*
 * sld_oracle.h - CPU restatement of the SPILADY SLDHL hot path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker for libspilady_hip.so: only
 * tests/, the smoke check and bench.py's cpu_baseline leg may load it. It is
 * never part of the product path.
 *
 * It follows the reference's CPU/OpenMP path (the *_CPU.cpp files) loop by
 * loop, with the same arithmetic order, single threaded, so that it can be
 * compared bit for bit with the compiled reference run on one OpenMP thread.
 * The reference's compile-time switches (control.h) are runtime fields of
 * sld_params. Noise comes either from the reference's MWC + Box-Muller
 * stream (random_number_CPU.cpp, consumed in loop order) or from the
 * counter-based Philox4x32-10 stream that the GPU engine uses (so GPU and
 * oracle agree step by step with thermostats on).
 */
#ifndef SLD_ORACLE_H
#define SLD_ORACLE_H

#include "../include/spilady_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_RNG_MWC 0
#define ORC_RNG_PHILOX 1

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const sld_params* p, const sld_atoms* init, int compute_initial_forces,
                    int rng_mode);
void orc_destroy(orc_ctx* o);

void orc_links(orc_ctx* o);
void orc_calculate_rho(orc_ctx* o);
void orc_calculate_force_energy(orc_ctx* o);
void orc_core_ds(orc_ctx* o, double dt);
void orc_core_dp(orc_ctx* o, double dt);
void orc_core_dr(orc_ctx* o, double dt);
void orc_core_dTe(orc_ctx* o, double dt);
void orc_calculate_temperature(orc_ctx* o);
void orc_check(orc_ctx* o, int step, sld_obs* out);
void orc_core(orc_ctx* o, int step);
int orc_step(orc_ctx* o, int first_step, int nsteps, sld_obs* obs);
void orc_scale_step(orc_ctx* o);
void orc_set_external(orc_ctx* o, const double* hext, const double* fext);
double orc_get_step(const orc_ctx* o);
/* the Berendsen barostats (scale_stress_CPU.cpp, scale_pressure_CPU.cpp) and
 * scale() (scale.cpp:31-44): stress, then pressure, then changestep */
void orc_set_barostat(orc_ctx* o, const sld_barostat* b);
void orc_scale(orc_ctx* o, int step);
void orc_get_box(const orc_ctx* o, double* box);
/* the quantum thermostat's filter H (quantum_noise_CPU.cpp:83-113) into out
 * (may be NULL); returns its length, 0 without quantumnoise */
int orc_quantum_H(const orc_ctx* o, double* out);

void orc_download(const orc_ctx* o, sld_atoms* out);
void orc_upload(orc_ctx* o, const sld_atoms* in);
void orc_get_cells(const orc_ctx* o, sld_cells* out);
void orc_set_cells_te(orc_ctx* o, const double* te);
void orc_get_geometry(const orc_ctx* o, int* ncell3, int* ngroups, int* group_sizes);
int orc_get_groups(const orc_ctx* o, int* members);
int orc_get_cell_lists(const orc_ctx* o, int cap, int* counts, int* lists);
double orc_total_time(const orc_ctx* o);

/* table `which` (0..15: F F' f f' V V' J J' A B C D A' B' C' D') */
double orc_table(const orc_ctx* o, int which, double x);
double orc_table_raw(const orc_ctx* o, int which, int n);
double orc_analytic(int which, double x);

/* DD05 pair term with knot coefficients V[8] (DD05.cpp:64-164), and the
 * coefficients a[k] (of r^k) of its quintic bridge from the end values
 * {y1, y1', y1'', y2, y2', y2''} at 1.8 and 2.0 A (DD05.cpp:88-158). */
double orc_pair_v(double rij, const double V[8]);
void orc_bridge_coeffs(const double ends[6], double a[6]);
/* Effective field sum_j J s_j + Landau + Heisenberg correction of every atom
 * at the current spins (calculate_spin_CPU.cpp:111-140,201-274), no update. */
void orc_fields(orc_ctx* o, double* h);

/* ---- parametric potential model (sld_model_oracle.c) ----
 * theta[22]: a_k[6] (density), V_k[8] (pair), aa, bb (embedding), J_A, J_rc
 * (exchange), Landau A..D. Kinds of generator: */
#define ORCM_EMBED 0   /* F(rho) */
#define ORCM_DENS 1    /* f(r) */
#define ORCM_PAIR 2    /* V(r) */
#define ORCM_EXCH 3    /* J(r) */
#define ORCM_NPARAM 22
typedef struct orcm_tab orcm_tab;
void orcm_default_params(double* theta);
double orcm_gen(const double* theta, int kind, double x);
/* d gen / d theta_k at x (zero when gen does not depend on theta_k) */
double orcm_dgen(const double* theta, int kind, int k, double x);
/* tables of the model at theta (10^6 samples, 6-point derivatives, as
 * potential_CPU.cpp:61-170) and of the derivative of each generator */
orcm_tab* orcm_build(const double* theta, double rcut_max);
void orcm_free(orcm_tab* t);
double orcm_table_raw(const orcm_tab* t, int kind, int deriv, int n);
/* One configuration (orthorhombic box lengths L[3]): energy, forces and
 * fields, and grad[k] += gE dE/dtheta_k + sum_i gF_i . dF_i/dtheta_k +
 * gH_i . dH_i/dtheta_k, by forward-mode differentiation one parameter at a
 * time. Returns 0, or -1 for a box shorter than twice rcut_max. */
int orcm_eval_grad(const orcm_tab* t, const double L[3], int n, const double* r, const double* s,
                   double rcut_max, double rcut_pot, double rcut_mag, double gE, const double* gF,
                   const double* gH, double* E, double* F, double* H, double* grad);

/* Philox4x32-10 (Salmon et al., SC'11) of one 128-bit counter. */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Three N(0,1) draws of the counter-based noise stream for one atom and one
 * noise call (shared definition with the GPU engine). */
void orc_normal3(uint64_t seed, int replica, int atom, uint64_t call, double out[3]);

#ifdef __cplusplus
}
#endif

#endif
