// Device state and kernel declarations of the spin-lattice engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sld_device.h"
#include "sld_model.h"

namespace sld {

// One independent system of a batch: its geometry, thermostat temperature
// and noise stream, and where its atoms and link cells sit in the
// concatenated arrays. A replica's atom slots start on a 256-slot boundary and
// its cells on an 8-cell boundary, so that every block of the per-atom and
// per-cell reductions belongs to one replica; slots and cells in the padding
// have replica -1.
struct Rep {
    Box box;
    double temperature;   // k_B T, eV
    double box_volume;
    uint64_t seed;
    int replica;          // noise stream id
    int natom, atom0;     // atoms [atom0, atom0 + natom)
    int ncell, cell0;     // link cells [cell0, cell0 + ncell)
};

// Per-atom arrays are indexed by atom slot (replica atom0 + the caller's atom
// index); vectors are double4 with .w used as a payload (s.w = |s|, others 0).
struct Dev {
    int N, C;                  // atom slots and cells over all replicas, padding included
    const Rep* rep;            // [R]
    const int* arep;           // replica of each atom slot, -1 for padding [N]
    const int* crep;           // replica of each cell, -1 for padding [C]
    double4 *r, *p, *f, *s, *hh, *hl;
    double4* zn;               // Langevin normals of the current spin pass [N]
    double4 *hext, *fext;      // external field (eV) and force (eV/A) [N]; null when off
    uint64_t* pass_base;       // noise passes consumed before the current step [1]
    // quantum thermostat: noise history ring [2 Nf][3N] (slot e % 2Nf holds
    // history element e), current noise [3N], filter of each replica [R][2 Nf];
    // null when off
    double *qring, *qval;
    const double* qH;
    double *rho, *pe, *ke, *me, *me0, *vir, *sjs, *fp;
    double4* lan;              // Landau A, B, C, D at each atom's density [N] (kept current with rho)
    double* stress;            // [6][N]: 11 22 33 12 23 31
    double* lvol;              // localvol: each atom's local volume
    double4* m;                // magmom: the moments m = -g s, the state between spin passes
    int* cell;                 // link cell of each atom (new_cell_index)
    int* cell_prev;            // link cell before the last links()
    int* cslot;                // cell * kCellCap + position in the cell's list
    int *cl, *cn;              // cell lists [C][kCellCap] head->tail, counts [C]
    int *inc, *inc_n;          // atoms that moved into each cell this step, any order [C][kCellCap], counts [C]
                               // (zero between steps: the relink that consumes them clears them)
    int *nb, *nbn;             // neighbours within rcut_max [N][kMaxNb]
    int *nbm, *nbmn;           // neighbours within rcut_mag [N][kMaxMag]
    double* nbJ;               // J(r) of the rcut_mag neighbours [N][kMaxMag]
    // electron subsystem, per link cell
    double *Te, *Ee, *Tl, *Gel, *Ges, *TsR, *TsL, *sRup, *sRdn, *sLup, *sLdn;
    double4* cellv;            // per cell: x,y,z = mean p (localcolmot)
    int* err;                  // error flag bits
};

// The electron mesh as the core_dTe sub-step kernels read it: only what they
// use, since a 2M-atom step launches about 100 of them and 3,600 sweeps.
struct Mesh {
    double *Te, *Ee;
    const double *Gel, *Ges, *Tl, *TsL;
    int C, lattlang, spinlang;
};

// Parameters shared by every replica of a batch.
struct Phys {
    double atmass, gamma_L, gamma_S;
    double rcut_pot, rcut_mag, rcut_pot_sq, rcut_mag_sq, rcut_max_sq;
    int spinlang, lattlang, eltemp, localcolmot;
    int extfield, extforce;
    int quantumnoise, qm, qnf2;  // quantum thermostat: on, redraw period M, filter taps 2 Nf
    int localvol;                // stress over per-atom local volumes
    int magmom;                  // moments as the spin variable
    double rcut_vol;             // their cut-off, A
};

// per-replica block of reduced scalars
constexpr int kScal = 64;
constexpr int kScalLvol = 56;  // localvol: the replica's sum of local volumes
// lanes per link cell in the momentum and temperature passes (a lane takes
// the cell's list entries l, l + kSeg, ...)
constexpr int kSeg = 8;

// ---- kernels ----
__global__ void k_bin(Dev D);
__global__ void k_relink(const int* cl_old, const int* cn_old, Dev D);
__global__ void k_nbuild(Tables T, Phys P, Dev D, int write_rho);
__global__ void k_fp(Tables T, Dev D);
__global__ void k_force(Tables T, Phys P, Dev D);
__global__ void k_lvol_sum(Dev D, double* partials);
__global__ void k_lvol_apply(Dev D, const double* scal);
__global__ void k_magmom_s(Dev D, int a0, int n);
__global__ void k_magmom_m(Dev D);
__global__ void k_spin_noise(Dev D, int pass_off);
__global__ void k_landau(Tables T, Dev D);
__global__ void k_step_max(Dev D, double* partials);
__global__ void k_scale_r(Dev D, double factor);
__global__ void k_rebox_r(Dev D, Box old_box, Box new_box);
__global__ void k_reduce_max2(const double* partials, int nblk, double* out);
__global__ void k_pass_advance(Dev D, int n);
__global__ void k_qn_init(Phys P, Dev D);
__global__ void k_qn_refresh(Phys P, Dev D, uint64_t r);
// LPC = 1, 2 or 4 lanes per cell; store_fields: write Heff_H, Heff_L and
// sum J|s_j| (the backward sweep of a core_ds call)
template <int LPC>
__global__ void k_spin_phase(const int* cells, int ncell, int backward, int store_fields, double dt, Tables T,
                             Phys P, Dev D);
__global__ void k_dp_cells(double dt, Phys P, Dev D, int pass_off, double* partials);
__global__ void k_dp_center(Phys P, Dev D, const double* scal, double* partials);
__global__ void k_dp_scale(Phys P, Dev D, const double* scal);
__global__ void k_dr(double dt, Phys P, Dev D);
__global__ void k_temperature(Tables T, Phys P, Dev D);
__global__ void k_dte_ab(double h, Mesh M, double* partials);
// the arguments are the few the stencil reads: one of these launches per
// level and sub-step, 3,600 per 2M-atom step
// cells: the level's cells as packed coordinates ix | iy << 10 | iz << 20;
// idx2, idy2, idz2: reciprocals of the squared cell edges (the stencil
// multiplies by them where the reference divides)
__global__ void k_dte_sweep(const int* cells, int ncell, double pref, double idx2, double idy2,
                            double idz2, int nx, int ny, int nz, double* Te, double* Ee, double* partials);
__global__ void k_reduce2(const double* pa, int na, const double* pb, int nb, double* out);
// the trailing small levels of a C sweep, run by one workgroup: level l of
// the run holds entries [off[l], off[l + 1]) of the sweep's level list, and
// its backward-sweep partials start at poff[l]
constexpr int kTailMaxLev = 32;
struct SweepTail {
    int nlev;
    int off[kTailMaxLev + 1];
    int poff[kTailMaxLev];
};
__global__ void k_dte_tail(const int* cells, SweepTail T, double pref, double idx2, double idy2, double idz2,
                           int nx, int ny, int nz, double* Te, double* Ee, double* partials, const double* pa,
                           int na, const double* pb, int nb, double* out);
__global__ void k_dte_finish(double h, Mesh M, const double* xp, int nxp);
__global__ void k_dte_finish_ab(double h, Mesh M, const double* xp, int nxp, double* partials);
__global__ void k_obs_atoms(Tables T, Phys P, Dev D, double* partials, int scan_spin);
__global__ void k_obs_cells(Dev D, double* partials);
__global__ void k_renorm(Dev D, const double* err, double* partials);
__global__ void k_reduce(const double* partials, int nblk, int nk, double* out);
__global__ void k_reduce_seg(const double* partials, const int* seg, int nk, double* out, int out_off);
__global__ void k_check_nonfinite(Dev D);
__global__ void k_spin_len(Dev D, int a0, int n);

// ---- parametric potential model (sld_model.hip) ----
struct ModelTab {
    const double* r;    // [N][kMRStride]
    const double* rd;   // [N][kMRDStride]
    const double* h;    // [N][kMHStride]
    double in_over_rmax, in_over_rhomax;
    double L[4];        // Landau A B C D
};
__global__ void k_model_rho(ModelTab M, Phys P, Dev D, double* rho_out, double* fp_out);
__global__ void k_model_forward(ModelTab M, Phys P, Dev D, const double* rho, const double* fp, const int* coff,
                                double* F, double* H, double* partials);
__global__ void k_model_q(ModelTab M, Phys P, Dev D, const int* coff, const double* gF, double* Q);
__global__ void k_model_adjoint(ModelTab M, Phys P, Dev D, const double* rho, const double* fp, const double* Q,
                                const int* coff, const double* gE, const double* gF, const double* gH,
                                double* partials);

constexpr int kObsAtomK = 17;  // pe ke me me0 sx sy sz rup rdn lup ldn, then (vir) stress[6]
constexpr int kObsCellK = 6;   // Ee sRup sRdn sLup sLdn Te

}  // namespace sld
