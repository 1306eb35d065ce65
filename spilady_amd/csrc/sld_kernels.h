// Device state and kernel declarations of the spin-lattice engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sld_device.h"

namespace sld {

// Per-atom arrays are indexed by the caller's atom index; vectors are double4
// with .w used as a payload (s.w = |s|, others 0).
struct Dev {
    int N, C;
    double4 *r, *p, *f, *s, *hh, *hl;
    double *rho, *pe, *ke, *me, *me0, *vir, *sjs, *fp;
    double* stress;            // [6][N]: 11 22 33 12 23 31
    int* cell;                 // link cell of each atom (new_cell_index)
    int* cell_prev;            // link cell before the last links()
    int* cslot;                // cell * kCellCap + position in the cell's list
    int *cl, *cn;              // cell lists [C][kCellCap] head->tail, counts [C]
    int *nb, *nbn;             // neighbours within rcut_max [N][kMaxNb]
    int *nbm, *nbmn;           // neighbours within rcut_mag [N][kMaxMag]
    double* nbJ;               // J(r) of the rcut_mag neighbours [N][kMaxMag]
    // electron subsystem, per link cell
    double *Te, *Ee, *Tl, *Gel, *Ges, *TsR, *TsL, *sRup, *sRdn, *sLup, *sLdn;
    double4* cellv;            // per cell: x,y,z = mean p (localcolmot)
    int* err;                  // error flag bits
};

struct Phys {
    double atmass, gamma_L, gamma_S, temperature;
    double rcut_pot, rcut_mag, rcut_pot_sq, rcut_mag_sq, rcut_max_sq;
    double box_volume;
    uint64_t seed;
    int replica;
    int spinlang, lattlang, eltemp, localcolmot;
};

// ---- kernels ----
__global__ void k_bin(Box b, Dev D);
__global__ void k_relink(Box b, const int* cl_old, const int* cn_old, Dev D);
__global__ void k_nbuild(Box b, Tables T, Phys P, Dev D, int write_rho);
__global__ void k_fp(Tables T, Dev D);
__global__ void k_force(Box b, Tables T, Phys P, Dev D);
__global__ void k_spin_phase(const int* cells, int ncell, int backward, double dt, Tables T, Phys P,
                             Dev D, uint64_t pass);
__global__ void k_dp_cells(double dt, Phys P, Dev D, uint64_t pass, double* partials);
__global__ void k_dp_center(Phys P, Dev D, const double* mean3, double* partials);
__global__ void k_dp_scale(Phys P, Dev D, const double* ke2);
__global__ void k_dr(double dt, Box b, Phys P, Dev D);
__global__ void k_temperature(Tables T, Phys P, Dev D);
__global__ void k_dte_ab(double h, Phys P, Dev D, double* partials);
__global__ void k_dte_sweep(const int* cells, int ncell, double pref, double dx2, double dy2,
                            double dz2, Box b, Dev D, double* partials);
__global__ void k_reduce2(const double* pa, int na, const double* pb, int nb, double* out);
__global__ void k_dte_finish(double h, Phys P, Dev D, const double* xp, int nxp);
__global__ void k_dte_finish_ab(double h, Phys P, Dev D, const double* xp, int nxp, double* partials);
__global__ void k_obs_atoms(Tables T, Phys P, Dev D, double* partials, int scan_spin);
__global__ void k_obs_cells(Dev D, double* partials);
__global__ void k_renorm(Dev D, const double* err, double* partials);
__global__ void k_reduce(const double* partials, int nblk, int nk, double* out);
__global__ void k_check_nonfinite(Dev D);

constexpr int kObsAtomK = 11;  // pe ke me me0 sx sy sz rup rdn lup ldn
constexpr int kObsCellK = 6;   // Ee sRup sRdn sLup sLdn Te

}  // namespace sld
