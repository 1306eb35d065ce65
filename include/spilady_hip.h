/*
 * spilady_hip.h - C ABI of libspilady_hip.so, the MI355X (gfx950) spin-lattice
 * dynamics engine for the SLDHL Hamiltonian: DD05 magnetic EAM plus Jij
 * Heisenberg exchange and Landau terms, FP64 throughout.
 *
 * The reference's operator API for this path is a set of argument-less,
 * global-state C++ functions selected at link time between *_CPU.cpp and
 * *_GPU.cu (prototype_CPU.h:120-289). Each entry point below replaces one of
 * them (cited per function); the state they shared through globals
 * (global.h:53-214) lives in an opaque context instead. The reference's
 * compile-time switches (control.h:8-43) are runtime fields of sld_params.
 *
 * Conventions:
 *  - every function returns 0 (SLD_OK) or a positive SLD_ERR_* code; the text
 *    of the last error is returned by sld_last_error(ctx);
 *  - host buffers are caller-owned and copied; device memory is owned by the
 *    context; vectors are [natom][3] row-major, stress is [natom][6] in the
 *    con- column order 11 22 33 12 23 31;
 *  - calls on one context are ordered on that context's HIP stream and are
 *    not re-entrant; one context per host thread.
 */
#ifndef SPILADY_HIP_H
#define SPILADY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLD_ABI_VERSION 2

#define SLD_OK 0
#define SLD_ERR_ARG 1       /* invalid argument or input state */
#define SLD_ERR_HIP 2       /* a HIP runtime call failed */
#define SLD_ERR_CAPACITY 3  /* a fixed per-cell/per-atom capacity was exceeded */
#define SLD_ERR_STATE 4     /* call not valid in the context's current state */
#define SLD_ERR_NOGPU 5     /* no usable gfx950 device */

/* Run parameters: the variables.in keys of read_variables.cpp:51-242 in
 * internal units (default.h:44-226), plus the control.h switches. */
typedef struct sld_params {
    int32_t natom;
    int32_t interval_of_print_out;  /* renormalizeEnergy acts on printed steps */
    double box[6];             /* d.xx d.yx d.yy d.zx d.zy d.zz (read_config.cpp:41-42) */
    double step;               /* time step, s */
    double temperature;        /* k_B*T, eV (read_variables.cpp:159) */
    double atmass;             /* eV/(A/s)^2 (read_variables.cpp:155-157) */
    double gamma_L_over_mass;  /* 1/s */
    double gamma_S_HL;         /* spin damping for the Heisenberg-Landau case */
    double kappa_e;            /* electron thermal conductivity, W/(m K) */
    double rcut_pot, rcut_mag, rcut_max;  /* A */
    double min_length_link_cell;          /* A */
    double start_time;         /* s */
    uint64_t random_seed;
    int32_t replica;           /* noise stream id (independent replicas) */
    int32_t spinlang;          /* control.h: spin Langevin thermostat */
    int32_t lattlang;          /* control.h: lattice Langevin thermostat */
    int32_t eltemp;            /* control.h: per-cell electron temperature */
    int32_t localcolmot;       /* control.h: local collective motion (needs eltemp) */
    int32_t renormalize_energy;/* control.h: renormalizeEnergy (needs eltemp) */
    int32_t changestep;        /* control.h: the time step adapts after every step (scale_step_CPU.cpp:33-97);
                                  needs a single replica */
    int32_t localvol;          /* control.h: each atom's stress over a local volume from its neighbours
                                  within rcut_vol (calculate_force_energy_CPU.cpp:80-90,266-283) instead of
                                  box_volume / natom */
    double displace_limit;     /* A, largest displacement per step (default.h:218) */
    double phi_limit;          /* rad, largest spin rotation per step (default.h:219) */
    int32_t extfield;          /* control.h: per-atom external field (external_field_CPU.cpp), zero
                                  until set with sld_set_external */
    int32_t extforce;          /* control.h: per-atom external force (external_force_CPU.cpp), same */
    int32_t quantumnoise;      /* control.h: quantum thermostat, coloured lattice noise
                                  (quantum_noise_CPU.cpp); needs lattlang */
    int32_t msteps_quantum;    /* the noise is redrawn every this many core_dp calls (default.h:223) */
    int32_t nfrequency_quantum;/* frequency intervals; the filter has 2x this many taps (default.h:224) */
    int32_t magmom;            /* control.h: magnetic moments m = -g s as the state's spin variable
                                  (core_ds_CPU.cpp:37-45,73-81; the J and Landau tables scaled by powers
                                  of g, potential_CPU.cpp:125-160) */
    double rcut_vol;           /* A, the localvol cut-off (default.h:208-211, read_variables.cpp:230) */
} sld_params;

/* Host view of the per-atom state (struct.h:44-120, con- columns of
 * write_config.cpp:92-131). Any pointer may be NULL in sld_download to skip
 * that field; r, p and s are required by sld_create. */
typedef struct sld_atoms {
    double *r, *p, *f, *s;     /* [natom][3]; with magmom s carries the moments m (read_config.cpp:96-99) */
    double *heff_h, *heff_l;
    double *rho, *pe, *ke, *me, *me0;
    double *stress;            /* [natom][6] */
    double *vir, *sum_jij_sj;
    double *local_volume;      /* A^3 (write_config.cpp:129); box_volume / natom without localvol */
} sld_atoms;

/* One printed step: the enr-, tmp- and spn- columns
 * (check_energy_CPU.cpp:121-133, check_temperature_CPU.cpp:144-158,
 * check_spin_CPU.cpp:66-76), the virial, and the str- and prs- columns
 * (check_stress_CPU.cpp:35-58, check_pressure_CPU.cpp:35-41) in GPa. */
typedef struct sld_obs {
    int32_t step;
    int32_t pad;
    double total_time;
    double ave_pe, ave_ke, ave_me, ave_me0, ave_ee;
    double tl, ts_r, ts_l, te;
    double ave_s[3], ave_s0, ave_m[3], ave_m0;
    double virial;
    double ave_stress[6];      /* 11 22 33 12 23 31 */
    double pressure;
} sld_obs;

/* Per-cell electron subsystem (struct.h:122-166); any pointer may be NULL. */
typedef struct sld_cells {
    double *te, *ee, *tl, *ts_r, *ts_l, *gel, *ges;
    double *sum_r_up, *sum_r_dn, *sum_l_up, *sum_l_dn;
} sld_cells;

typedef struct sld_ctx sld_ctx;

/* Library identification: ABI version and the device it bound to. */
int sld_version(void);
/* Host-only planning, no device call: link-cell counts for a box
 * (map_cells.cpp:41-60), the greedy cell groups (allocate_cells_CPU.cpp:45-122)
 * and the dependency level of every cell in the four group-ordered sweeps the
 * engine launches level by level (spin forward/backward over the 26
 * neighbours, electron diffusion forward/backward over the 6 faces).
 * Call with group_sizes == NULL to get ncell3 only. Arrays hold C =
 * ncell3[0]*ncell3[1]*ncell3[2] entries (levels: 4*C, schedule-major). */
int sld_plan(const double* box, double min_length_link_cell, int* ncell3, int* ngroups,
             int* group_sizes, int* group_cells, int* nlevels, int* levels);
int sld_device_info(char* name, int len, int* n_cu);

/* Replaces initial_GPU + potential_table_GPU + copy_CPU_to_GPU +
 * initial_links + allocate_cells + initialize_random_number and, with eltemp,
 * the map_cells Te set-up and the initial calculate_temperature
 * (initial.cpp:31-106, initial_GPU.cu:36-102, potential_GPU.cu:43-96,
 * copy_CPU_btw_GPU.cu:35-138, links_CPU.cpp:33-83,
 * allocate_cells_CPU.cpp:33-125, map_cells.cpp:128-159).
 * With readconf the reference takes f and rho from the input (initial.cpp:78-83);
 * compute_initial_forces = 1 evaluates them instead (the non-readconf path). */
int sld_create(const sld_params* params, const sld_atoms* init,
               int compute_initial_forces, sld_ctx** out);
/* Batched replicas: nrep independent systems (params[r], init[r]) of any
 * size, temperature and seed in one context; every step entry point advances
 * all of them, each kernel launch covering all replicas. No counterpart in the
 * reference, which runs one system per process; each replica follows the same
 * reference functions as sld_create's single system, and its trajectory is
 * bitwise the one sld_create gives it alone (noise is keyed by
 * (random_seed, replica) and the replica's own atom index). The replicas must
 * share dt, mass, damping, cutoffs, print interval and switches; eltemp needs
 * nrep == 1. sld_create(p, a, f, out) is sld_create_batch(1, p, a, f, out). */
int sld_create_batch(int nrep, const sld_params* params, const sld_atoms* init,
                     int compute_initial_forces, sld_ctx** out);
int sld_nreplicas(const sld_ctx* ctx);
/* The replica that sld_download, sld_upload, sld_get_cells, sld_set_cells_te,
 * sld_get_geometry and sld_get_cell_lists act on (default 0). */
int sld_select_replica(sld_ctx* ctx, int replica);
/* Replaces free_memory (free_memory.cpp:31-52). */
int sld_destroy(sld_ctx* ctx);
const char* sld_last_error(const sld_ctx* ctx);

/* links (links_CPU.cpp:85-177 / links_GPU.cu:67-285). */
int sld_links(sld_ctx* ctx);
/* calculate_rho (calculate_rho_CPU.cpp:71-97). */
int sld_calculate_rho(sld_ctx* ctx);
/* calculate_force_energy (calculate_force_energy_CPU.cpp:220-301). */
int sld_calculate_force_energy(sld_ctx* ctx);
/* core_ds (core_ds_CPU.cpp:35-83). */
int sld_core_ds(sld_ctx* ctx, double dt);
/* core_dp (core_dp_CPU.cpp:51-85). */
int sld_core_dp(sld_ctx* ctx, double dt);
/* core_dr (core_dr_CPU.cpp:33-44). */
int sld_core_dr(sld_ctx* ctx, double dt);
/* core_dTe (core_dTe_CPU.cpp:47-64); needs eltemp. */
int sld_core_dTe(sld_ctx* ctx, double dt);
/* calculate_temperature (calculate_temperature_CPU.cpp:35-148); needs eltemp. */
int sld_calculate_temperature(sld_ctx* ctx);
/* check (check.cpp:31-47): check_energy (with the renormalizeEnergy feedback,
 * check_energy_CPU.cpp:77-111), check_temperature, check_spin. step = -1
 * records the initial energy. Writes one record per replica to out[0..nrep)
 * (may be NULL). */
int sld_check(sld_ctx* ctx, int step, sld_obs* out);
/* nsteps of core() (core.cpp:31-76), each followed by total_time += step and
 * check() on printed steps, as the driver loop does (spilady.cpp:50-57).
 * obs (may be NULL) receives nrep records per printed step, printed-step
 * major (obs[k * nrep + r]); *nobs (may be NULL) the number of printed steps. */
int sld_step(sld_ctx* ctx, int first_step, int nsteps, sld_obs* obs, int* nobs);
/* scale_step (scale_step_CPU.cpp:33-97, called through scale(), scale.cpp): the
 * largest displacement |p|max/m dt and spin rotation |H|max/hbar dt of the
 * step just taken; dt *= 0.8 if either exceeds its limit, else dt *= 1.05.
 * sld_step calls it after every step when changestep is on. */
int sld_scale_step(sld_ctx* ctx);
/* The current time step (s). */
double sld_get_step(const sld_ctx* ctx);

/* Berendsen-type barostats, the control.h STRESS and PRESSURE options
 * (scale_stress_CPU.cpp:33-86, scale_pressure_CPU.cpp:33-61). Off until set;
 * single-system contexts only. The box changes, the link-cell counts do not
 * (as in the reference). */
typedef struct sld_barostat {
    int32_t stress_on;                  /* STRESS: box matrix toward the target stresses */
    int32_t pressure_on;                /* PRESSURE: isotropic, toward the target pressure */
    int32_t interval_of_scale_stress;   /* scale after every this many steps */
    int32_t interval_of_scale_pressure;
    double stress[6];                   /* GPa: stress_xx yx yy zx zy zz (read_variables.cpp:197-203) */
    double pressure;                    /* GPa (read_variables.cpp:209) */
    double baro_damping_time;           /* s (read_variables.cpp:213) */
} sld_barostat;
int sld_set_barostat(sld_ctx* ctx, const sld_barostat* b);
/* scale(step) (scale.cpp:31-44): STRESS, then PRESSURE at their intervals,
 * then changestep. It uses the stresses and pressure of the last check().
 * sld_step calls it after every step itself. */
int sld_scale(sld_ctx* ctx, int step);
/* The two barostats one at a time, for callers that run scale() themselves:
 * scale_stress (scale_stress_CPU.cpp:33-86) and scale_pressure
 * (scale_pressure_CPU.cpp:33-61). Each acts unconditionally (the caller
 * keeps the interval test of scale.cpp:33-39) on the stresses or pressure of
 * the last sld_check, over the time since that barostat last acted. */
int sld_scale_stress(sld_ctx* ctx);
int sld_scale_pressure(sld_ctx* ctx);
/* Sets the context's clock, the reference's global total_time (global.h:55),
 * for callers that advance it themselves (spilady.cpp:52): sld_check reports
 * it and the barostats measure their intervals on it. sld_step advances the
 * clock by itself. */
int sld_set_total_time(sld_ctx* ctx, double total_time);
/* The current box matrix of the selected replica (d.xx d.yx d.yy d.zx d.zy d.zz). */
int sld_get_box(sld_ctx* ctx, double* box);
/* The per-atom external field Hext (eV, i.e. -g muB B; external_field_CPU.cpp:
 * 38-53) and external force fext (eV/A; external_force_CPU.cpp:38-50) of the
 * selected replica, [natom][3] each in atom order; NULL leaves one unchanged.
 * Needs the matching switch in sld_params. Hext starts every spin field sum
 * (calculate_spin_CPU.cpp:35-39) and enters me as -s.Hext
 * (calculate_force_energy_CPU.cpp:199-201); core_dp adds fext to f in place
 * before the momentum update (core_dp_CPU.cpp:69-71,130-132). The reference
 * reads the field before its initial calculate_temperature (initial.cpp:58-
 * 106): with eltemp, call sld_calculate_temperature after setting it at the
 * start of a run. */
int sld_set_external(sld_ctx* ctx, const double* hext, const double* fext);

/* Device <-> host state of the selected replica (copy_atoms_from_GPU_to_CPU,
 * copy_CPU_btw_GPU.cu:141-150). Atoms are in the caller's original index order. */
int sld_download(sld_ctx* ctx, sld_atoms* out);
/* Overwrites r, p, s (and f, rho when non-NULL) and rebuilds nothing else:
 * the cell lists keep their current order. */
int sld_upload(sld_ctx* ctx, const sld_atoms* in);
int sld_get_cells(sld_ctx* ctx, sld_cells* out);
/* Te (k_B T, eV) of every cell of the selected replica, in cell order, and
 * Ee = E(Te) with it: the readTe start (map_cells.cpp:136-157). */
int sld_set_cells_te(sld_ctx* ctx, const double* te);
/* Link-cell geometry and groups (map_cells.cpp:41-60, allocate_cells_CPU.cpp). */
int sld_get_geometry(sld_ctx* ctx, int* ncell3, int* ngroups, int* group_sizes);
/* Per-cell linked lists, head to tail, as atom indices; cap entries per cell. */
int sld_get_cell_lists(sld_ctx* ctx, int cap, int* counts, int* lists);
double sld_total_time(const sld_ctx* ctx);

/* Device time of the work issued by the last sld_step call, measured with
 * HIP events on the context's stream. */
int sld_last_step_timing(sld_ctx* ctx, double* total_ms);
/* Bracket every launch of the kernel named `kernel` (e.g. "k_force"; "" or
 * NULL turns it off) with HIP events on the context's stream; after each
 * sld_step, sld_kernel_timing returns the launches seen and their summed
 * device time over that call. */
int sld_time_kernel(sld_ctx* ctx, const char* kernel);
int sld_kernel_timing(sld_ctx* ctx, int* launches, double* total_ms);

/* ---- Launch shapes and the autotuner ----
 * No counterpart in the reference (its CUDA launches use fixed block sizes).
 * These are the block shapes of the kernels whose results do not depend on
 * them: every reduction inside them sums per wavefront, so a trajectory is
 * bitwise the same under any valid shapes. */
typedef struct {
    int spin_block;        /* core_ds sweep level: threads (one per cell) per block; 32, 64 or 128 */
    int sweep_block;       /* core_dTe diffusion level: threads per block; 64, 128 or 256 */
    int force_block;       /* force/energy: threads (one per atom) per block; 64, 128 or 256 */
    int nbuild_cells;      /* neighbour build: link cells (one wave each) per block; 1, 2 or 4 */
    int temperature_block; /* calculate_temperature: threads (32 per cell) per block; 64, 128 or 256 */
    int sweep_tail;        /* core_dTe diffusion: the trailing levels of each sweep with at most this many
                              cells run in one workgroup, level after level (0: every level its own
                              launch); 0, 2048 or 8192 */
    int spin_lanes;        /* core_ds sweep level: lanes that share one cell's walk (the neighbour
                              gathers and field sums split among them); 1, 2 or 4 */
} sld_shapes;
int sld_get_shapes(sld_ctx* ctx, sld_shapes* out);
/* How sld_step issues its work (no counterpart in the reference): 0 (default)
 * eager launches; 1 core_dTe replayed from a captured graph; 2 each
 * core() step replayed from a captured graph (one per cell-list buffer
 * parity). Results are bitwise the same in every mode. The environment
 * variable SLD_GRAPHS=0|1|2 sets the mode at creation. */
int sld_set_graph_mode(sld_ctx* ctx, int mode);
int sld_set_shapes(sld_ctx* ctx, const sld_shapes* in);
/* Times each candidate shape of each kernel over `steps` full steps (after
 * one warm-up step) with HIP events, keeps the fastest, and restores the
 * context's state afterwards, so the trajectory continues exactly as if the
 * call had not been made. best (may be NULL) receives the chosen shapes;
 * us (may be NULL, 21 entries: kernel-major in the order of sld_shapes, three
 * candidates each in ascending order) the device microseconds per step of
 * each candidate, 0 where the kernel did not run. Needs device memory for a
 * copy of the context's state while it runs. */
int sld_autotune(sld_ctx* ctx, int steps, sld_shapes* best, double* us);

/* ---- Parametric potential model and its parameter adjoint ----
 * No counterpart in the reference, whose potential (DD05.cpp, JijFe.cpp) is
 * fixed at compile time; these are for callers that fit the potential's
 * coefficients to energies, forces and effective fields. The configurations
 * are the context's replicas, frozen at their current positions and spins.
 * The parameter vector theta has sld_model_nparams() = 22 entries:
 *   [0..5]   a_k of the density f(r) = sum a_k (R_k - r)^3          (DD05.cpp:20-43)
 *   [6..13]  V_k of the pair term V(r) = sum V_k (R_k - r)^3, r > 2  (DD05.cpp:64-164;
 *            the ZBL core below 1.8 and the quintic bridge follow)
 *   [14..15] aa, bb of the embedding F(rho)                          (DD05.cpp:6-18)
 *   [16..17] A, rc of the exchange J(r) = A (1 - r/rc)^3              (JijFe.cpp:5-14)
 *   [18..21] Landau A, B, C, D                                        (JijFe.cpp:18-43)
 * E_r = sum over replica r's atoms of pe + me + me0; F = -dE/dr and
 * H = -dE/ds per atom. Atom arrays are compact over the replicas, replica
 * r's atoms after those of replicas 0..r-1, each in its own index order
 * (sld_model_natoms atoms in all). Buffers may be host or device memory. */
int sld_model_nparams(void);
/* The reference's values of theta. */
int sld_model_default_params(double* theta);
/* Builds the model's tables (10^6 samples each, as potential_CPU.cpp:61-170)
 * at theta; about a second on the host. */
int sld_model_set_params(sld_ctx* ctx, const double* theta);
int sld_model_get_params(sld_ctx* ctx, double* theta);
int sld_model_natoms(const sld_ctx* ctx);
/* energy[nrep], force[3 * natoms], field[3 * natoms]; any may be NULL. */
int sld_model_eval(sld_ctx* ctx, double* energy, double* force, double* field);
/* grad[22] = sum_r g_energy[r] dE_r/dtheta + sum_i g_force_i . dF_i/dtheta
 *          + sum_i g_field_i . dH_i/dtheta, exact for the tabulated model;
 * NULL upstream gradients count as zero. */
int sld_model_vjp(sld_ctx* ctx, const double* g_energy, const double* g_force, const double* g_field,
                  double* grad);
/* sld_model_eval for callers that cannot allocate device memory (code under
 * stream capture, serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference. force and field are device buffers (or
 * NULL), written directly; energy may be host or device. It allocates
 * nothing: its scratch (density and F'(rho) per atom slot, energy partials)
 * is a fixed array in the library's device image, 4.2 MB on each GPU it runs
 * on, held until the process ends. Contexts of more than 262,144 atom slots
 * (padding included) or 64 replicas are refused with SLD_ERR_CAPACITY. All
 * calls on one GPU share that scratch: calls that overlap on one GPU (other
 * streams or host threads) race. Results are bitwise those of
 * sld_model_eval. */
int sld_model_eval_noalloc(sld_ctx* ctx, double* energy, double* force, double* field);

#ifdef __cplusplus
}
#endif

#endif /* SPILADY_HIP_H */
