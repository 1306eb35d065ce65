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

/* Philox4x32-10 (Salmon et al., SC'11) of one 128-bit counter. */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Three N(0,1) draws of the counter-based noise stream for one atom and one
 * noise call (shared definition with the GPU engine). */
void orc_normal3(uint64_t seed, int replica, int atom, uint64_t call, double out[3]);

#ifdef __cplusplus
}
#endif

#endif
