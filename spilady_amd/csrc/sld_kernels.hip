// HIP kernels of the spin-lattice engine (gfx950, FP64).
//
// Work decomposition: one thread per atom for the pair passes (neighbour
// build + density, force), one thread per link cell for everything that the
// reference runs as a sequential walk down a cell's list (spin sweeps,
// Langevin momentum with local collective motion, cell temperatures).
// Group-ordered sweeps are launched one dependency level at a time by the
// host; within a launch no two cells interact.
#include <hip/hip_runtime.h>

#include "../../include/spilady_hip.h"
#include "sld_kernels.h"

namespace sld {

// ---------------------------------------------------------------------------
// block reductions into per-block partials (fixed tree order: deterministic)

template <int NK>
__device__ __forceinline__ void block_partials(const double (&v)[NK], double* partials) {
    __shared__ double sh[NK][256];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NK; ++k) sh[k][t] = v[k];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int k = 0; k < NK; ++k) sh[k][t] += sh[k][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < NK; ++k) partials[(size_t)blockIdx.x * NK + k] = sh[k][0];
    }
}

// out[k] = sum over blocks of partials[b*nk + k]; one block of 256 threads.
__global__ __launch_bounds__(256) void k_reduce(const double* partials, int nblk, int nk, double* out) {
    __shared__ double sh[256];
    const int t = threadIdx.x;
    for (int k = 0; k < nk; ++k) {
        double acc = 0.0;
        for (int b = t; b < nblk; b += 256) acc += partials[(size_t)b * nk + k];
        sh[t] = acc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) sh[t] += sh[t + s];
            __syncthreads();
        }
        if (t == 0) out[k] = sh[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// links (links_CPU.cpp:85-177)

__device__ __forceinline__ bool cells_adjacent(const Box& b, int c0, int c1) {
    int x0 = c0 % b.nx, y0 = (c0 / b.nx) % b.ny, z0 = c0 / (b.nx * b.ny);
    int x1 = c1 % b.nx, y1 = (c1 / b.nx) % b.ny, z1 = c1 / (b.nx * b.ny);
    int dx = (x1 - x0 + b.nx) % b.nx, dy = (y1 - y0 + b.ny) % b.ny, dz = (z1 - z0 + b.nz) % b.nz;
    return (dx <= 1 || dx == b.nx - 1) && (dy <= 1 || dy == b.ny - 1) && (dz <= 1 || dz == b.nz - 1);
}

// new cell of every atom (clamped binning of links_CPU.cpp:96-112)
__global__ __launch_bounds__(256) void k_bin(Box b, Dev D) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    int old = D.cell[a];
    int c = bin_cell(b, d3(D.r[a]));
    D.cell_prev[a] = old;
    D.cell[a] = c;
    if (c != old && !cells_adjacent(b, old, c)) atomicOr(D.err, kErrFarMove);
}

// Rebuild cell c's list. The reference detaches every mover and pushes it at
// the head of its new cell in ascending atom index (links_CPU.cpp:144-173),
// so a cell's new list is [movers in, descending index] + [stayers, old order].
__global__ __launch_bounds__(128) void k_relink(Box b, const int* cl_old, const int* cn_old, Dev D) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.C) return;
    int mv[kCellCap];
    int nm = 0;
    bool over = false;
    for (int slot = 0; slot < 26; ++slot) {
        int nc = neigh_cell(b, c, slot);
        int n = cn_old[nc];
        for (int k = 0; k < n; ++k) {
            int a = cl_old[nc * kCellCap + k];
            if (D.cell[a] == c) {
                if (nm < kCellCap) {
                    // insertion keeping descending atom index
                    int q = nm++;
                    while (q > 0 && mv[q - 1] < a) { mv[q] = mv[q - 1]; --q; }
                    mv[q] = a;
                } else {
                    over = true;
                }
            }
        }
    }
    int* out = D.cl + c * kCellCap;
    int k = 0;
    for (int q = 0; q < nm; ++q) out[k++] = mv[q];
    int n = cn_old[c];
    for (int q = 0; q < n; ++q) {
        int a = cl_old[c * kCellCap + q];
        if (D.cell[a] == c) {
            if (k < kCellCap) out[k++] = a;
            else over = true;
        }
    }
    D.cn[c] = k;
    if (over) atomicOr(D.err, kErrCellOverflow);
}

// ---------------------------------------------------------------------------
// neighbour list + density (calculate_rho_CPU.cpp:33-92 for rho, the 27-cell
// scan order of calculate_spin_CPU.cpp:201-235 for the lists)

__global__ __launch_bounds__(256) void k_nbuild(Box b, Tables T, Phys P, Dev D, int write_rho) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.N) return;
    const int ci = D.cell[i];
    const double3 ri = d3(D.r[i]);
    int nn = 0, nm = 0;
    double rho = 0.0;
    bool over = false;
    int* nb = D.nb + (size_t)i * kMaxNb;
    int* nbm = D.nbm + (size_t)i * kMaxMag;
    double* nbJ = D.nbJ + (size_t)i * kMaxMag;
    for (int slot = 0; slot <= 26; ++slot) {
        int wc = slot < 26 ? neigh_cell(b, ci, slot) : ci;
        int n = D.cn[wc];
        const int* lst = D.cl + wc * kCellCap;
        for (int k = 0; k < n; ++k) {
            int j = lst[k];
            if (j == i) continue;
            double3 rij = min_image(b, sub3(ri, d3(D.r[j])));
            double rsq = sq3(rij);
            if (rsq < P.rcut_max_sq) {
                if (nn < kMaxNb) nb[nn++] = j; else over = true;
                double r = sqrt(rsq);
                int n0;
                double ratio;
                interp_index(r, T.in_over_rmax, n0, ratio);
                if (r < P.rcut_pot) rho += lerp_tab(T.rtab, kRStride, R_SF, n0, ratio);
                if (rsq < P.rcut_mag_sq) {
                    if (nm < kMaxMag) {
                        nbm[nm] = j;
                        nbJ[nm] = lerp_tab(T.rtab, kRStride, R_J, n0, ratio);
                        ++nm;
                    } else {
                        over = true;
                    }
                }
            }
        }
    }
    D.nbn[i] = nn;
    D.nbmn[i] = nm;
    if (write_rho) D.rho[i] = rho;
    if (over) atomicOr(D.err, kErrNbOverflow);
}

// ---------------------------------------------------------------------------
// force and energy (calculate_force_energy_CPU.cpp:33-301), full neighbour
// list: each atom sums its own pair terms; no atomics.

// dF/drho plus the Landau derivative terms of one atom
// (calculate_force_energy_CPU.cpp:98-107)
__global__ __launch_bounds__(256) void k_fp(Tables T, Dev D) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.N) return;
    int n0;
    double ratio;
    interp_index(D.rho[i], T.in_over_rhomax, n0, ratio);
    double s2 = sq3(d3(D.s[i]));
    double dF = lerp_tab(T.rhotab, kRhoStride, H_DF, n0, ratio);
    double dA = lerp_tab(T.rhotab, kRhoStride, H_DA, n0, ratio);
    double dB = lerp_tab(T.rhotab, kRhoStride, H_DB, n0, ratio);
    double dC = lerp_tab(T.rhotab, kRhoStride, H_DC, n0, ratio);
    double dD = lerp_tab(T.rhotab, kRhoStride, H_DD, n0, ratio);
    double s4 = s2 * s2;
    D.fp[i] = dF + (dA * s2 + dB * s4 + dC * s4 * s2 + dD * s4 * s4);
}

__global__ __launch_bounds__(256) void k_force(Box b, Tables T, Phys P, Dev D) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.N) return;
    const double3 ri = d3(D.r[i]);
    const double4 si4 = D.s[i];
    const double3 si = d3(si4);
    const double si_sq = sq3(si);
    const double fpi = D.fp[i];
    double3 f = make_double3(0.0, 0.0, 0.0);
    double pe = 0.0, me = 0.0, me0 = 0.0, vir = 0.0;
    double st0 = 0.0, st1 = 0.0, st2 = 0.0, st3 = 0.0, st4 = 0.0, st5 = 0.0;
    const int nn = D.nbn[i];
    const int* nb = D.nb + (size_t)i * kMaxNb;
    for (int k = 0; k < nn; ++k) {
        int j = nb[k];
        double3 rij = min_image(b, sub3(ri, d3(D.r[j])));
        double rsq = sq3(rij);
        double r = sqrt(rsq);
        int n0;
        double ratio;
        interp_index(r, T.in_over_rmax, n0, ratio);
        const double* t0 = T.rtab + (size_t)(n0 - 1) * kRStride;
        const double* t1 = t0 + kRStride;
        const double wr = 1.0 - ratio;
        double dudr = 0.0;
        if (r < P.rcut_pot) {
            double dsf = t0[R_DSF] * ratio + t1[R_DSF] * wr;
            double dpr = t0[R_DPR] * ratio + t1[R_DPR] * wr;
            double pr = t0[R_PR] * ratio + t1[R_PR] * wr;
            dudr = (fpi + D.fp[j]) * dsf + dpr;
            pe += 0.5 * pr;
        }
        if (r < P.rcut_mag) {
            double3 sj = d3(D.s[j]);
            double sdot = dot3(si, sj);
            double stimes = sqrt(si_sq * sq3(sj));
            double J = t0[R_J] * ratio + t1[R_J] * wr;
            double dJ = t0[R_DJ] * ratio + t1[R_DJ] * wr;
            dudr += -dJ * (sdot - stimes);
            double Jh = J / 2.0;
            me0 += Jh * stimes;
            me += -Jh * sdot;
        }
        double force = -dudr / r;
        double3 fij = mul3(force, rij);
        f = add3(f, fij);
        st0 += fij.x * rij.x; st1 += fij.y * rij.y; st2 += fij.z * rij.z;
        st3 += fij.x * rij.y; st4 += fij.y * rij.z; st5 += fij.z * rij.x;
        vir += -force * rsq;
    }
    // one-body terms (calculate_force_energy_CPU.cpp:196-215)
    int n0;
    double ratio;
    interp_index(D.rho[i], T.in_over_rhomax, n0, ratio);
    double F = lerp_tab(T.rhotab, kRhoStride, H_F, n0, ratio);
    double A = lerp_tab(T.rhotab, kRhoStride, H_A, n0, ratio);
    double B = lerp_tab(T.rhotab, kRhoStride, H_B, n0, ratio);
    double C = lerp_tab(T.rhotab, kRhoStride, H_C, n0, ratio);
    double Dd = lerp_tab(T.rhotab, kRhoStride, H_D, n0, ratio);
    double s4 = si_sq * si_sq;
    pe += F;
    me += A * si_sq + B * s4 + C * s4 * si_sq + Dd * s4 * s4;
    D.f[i] = make_double4(f.x, f.y, f.z, 0.0);
    D.pe[i] = pe;
    D.me[i] = me;
    D.me0[i] = me0;
    // the reference books -force*r^2 once per pair on the half-shell owner;
    // here each atom keeps half of each of its pairs (same total virial)
    D.vir[i] = 0.5 * vir;
    // per-atom stress with the kinetic part (calculate_force_energy_CPU.cpp:283-293)
    const double lv = P.box_volume / D.N;
    const double m = P.atmass;
    const double4 p = D.p[i];
    const size_t N = D.N;
    D.stress[0 * N + i] = (p.x * p.x / m + st0 / 2.0) / lv;
    D.stress[1 * N + i] = (p.y * p.y / m + st1 / 2.0) / lv;
    D.stress[2 * N + i] = (p.z * p.z / m + st2 / 2.0) / lv;
    D.stress[3 * N + i] = ((p.x * p.y) / m + st3 / 2.0) / lv;
    D.stress[4 * N + i] = ((p.y * p.z) / m + st4 / 2.0) / lv;
    D.stress[5 * N + i] = ((p.z * p.x) / m + st5 / 2.0) / lv;
}

// ---------------------------------------------------------------------------
// spins (calculate_spin_CPU.cpp:33-199, core_ds_CPU.cpp:35-127)

// rotation of s about Heff by |Heff| dt / hbar (core_ds_CPU.cpp:90-127)
__device__ __forceinline__ double3 spin_rotation(double3 H, double3 s, double dt) {
    double3 w = make_double3(H.x / -kHbar, H.y / -kHbar, H.z / -kHbar);
    double w0 = sqrt(sq3(w));
    if (w0 > 0.0) w = make_double3(w.x / w0, w.y / w0, w.z / w0);
    else w = make_double3(0.0, 0.0, 0.0);
    double o12 = w.x * w.y, o23 = w.y * w.z, o13 = w.x * w.z;
    double o1s = w.x * w.x, o2s = w.y * w.y, o3s = w.z * w.z;
    double A, cw;
    sincos(w0 * dt, &A, &cw);
    double B = 1.0 - cw;
    double3 t;
    t.x = s.x + (s.x * B * (-o2s - o3s) + s.y * (B * o12 - A * w.z) + s.z * (A * w.y + B * o13));
    t.y = s.y + (s.y * B * (-o1s - o3s) + s.z * (B * o23 - A * w.x) + s.x * (A * w.z + B * o12));
    t.z = s.z + (s.z * B * (-o1s - o2s) + s.x * (B * o13 - A * w.y) + s.y * (A * w.x + B * o23));
    return t;
}

// Landau field plus the Heisenberg correction -(sum J|s_j|/|s|) s
// (calculate_spin_CPU.cpp:131-140)
__device__ __forceinline__ double3 landau_field(double3 s, double A, double B, double C, double Dd,
                                                double sjs) {
    double s2 = sq3(s);
    double s4 = s2 * s2;
    double k = -(2.0 * A + 4.0 * B * s2 + 6.0 * C * s4 + 8.0 * Dd * s4 * s2);
    double s0 = sqrt(s2);
    double hc = s0 > 0.0 ? -sjs / s0 : 0.0;
    return make_double3((k + hc) * s.x, (k + hc) * s.y, (k + hc) * s.z);
}

__device__ __forceinline__ void spin_update(int a, int c, double dt, const Tables& T, const Phys& P,
                                            const Dev& D, uint64_t pass) {
    double3 H = make_double3(0.0, 0.0, 0.0);
    double S = 0.0;
    const int nm = D.nbmn[a];
    const int* nbm = D.nbm + (size_t)a * kMaxMag;
    const double* nbJ = D.nbJ + (size_t)a * kMaxMag;
    for (int k = 0; k < nm; ++k) {
        double J = nbJ[k];
        double4 sj = D.s[nbm[k]];
        H.x += J * sj.x; H.y += J * sj.y; H.z += J * sj.z;
        S += J * sj.w;
    }
    double3 s = d3(D.s[a]);
    double3 L = make_double3(0.0, 0.0, 0.0);
    if (!P.spinlang) {
        s = spin_rotation(H, s, dt);
    } else {
        const double dth = dt / 2.0, dtq = dt / 4.0, g = P.gamma_S;
        s = spin_rotation(H, s, dth);
        int n0;
        double ratio;
        interp_index(D.rho[a], T.in_over_rhomax, n0, ratio);
        double A = lerp_tab(T.rhotab, kRhoStride, H_A, n0, ratio);
        double B = lerp_tab(T.rhotab, kRhoStride, H_B, n0, ratio);
        double C = lerp_tab(T.rhotab, kRhoStride, H_C, n0, ratio);
        double Dd = lerp_tab(T.rhotab, kRhoStride, H_D, n0, ratio);
        // RK2
        L = landau_field(s, A, B, C, Dd, S);
        double3 st = add3(s, mul3(g * dtq, add3(H, L)));
        L = landau_field(st, A, B, C, Dd, S);
        s = add3(s, mul3(g * dth, add3(H, L)));
        // noise
        double Tk = P.eltemp ? D.Te[c] : P.temperature;
        double rs = sqrt(2.0 * Tk * g / dt);
        double3 z = normal3(P.seed, P.replica, a, pass);
        s = add3(s, mul3(dt, mul3(rs, z)));
        // RK2
        L = landau_field(s, A, B, C, Dd, S);
        st = add3(s, mul3(g * dtq, add3(H, L)));
        L = landau_field(st, A, B, C, Dd, S);
        s = add3(s, mul3(g * dth, add3(H, L)));
        s = spin_rotation(H, s, dth);
        D.sjs[a] = S;
    }
    D.s[a] = make_double4(s.x, s.y, s.z, sqrt(sq3(s)));
    D.hh[a] = make_double4(H.x, H.y, H.z, 0.0);
    D.hl[a] = make_double4(L.x, L.y, L.z, 0.0);
}

// One dependency level of a core_ds sweep: every listed cell walks its list
// head to tail (forward) or tail to head (backward).
__global__ __launch_bounds__(64) void k_spin_phase(const int* cells, int ncell, int backward, double dt,
                                                   Tables T, Phys P, Dev D, uint64_t pass) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    int c = cells[t];
    int n = D.cn[c];
    const int* lst = D.cl + c * kCellCap;
    if (!backward) {
        for (int k = 0; k < n; ++k) spin_update(lst[k], c, dt, T, P, D, pass);
    } else {
        for (int k = n - 1; k >= 0; --k) spin_update(lst[k], c, dt, T, P, D, pass);
    }
}

// ---------------------------------------------------------------------------
// momenta (core_dp_CPU.cpp:51-312). Every part of the Langevin update is local
// to one link cell, so a thread runs C1, B, A, B, C2 for its cell in order.

__device__ __forceinline__ void dp_colmot(const int* lst, int n, double factor, bool backward, Dev D) {
    for (int q = 0; q < n; ++q) {
        int a = lst[backward ? n - 1 - q : q];
        double3 sp = make_double3(0.0, 0.0, 0.0);
        for (int w = 0; w < n; ++w) sp = add3(sp, d3(D.p[lst[w]]));
        double3 pa = d3(D.p[a]);
        pa = add3(pa, mul3(factor, sub3(sp, pa)));
        D.p[a] = make_double4(pa.x, pa.y, pa.z, 0.0);
    }
}

__global__ __launch_bounds__(256) void k_dp_cells(double dt, Phys P, Dev D, uint64_t pass,
                                                  double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (c < D.C) {
        const int n = D.cn[c];
        const int* lst = D.cl + c * kCellCap;
        const double gm = P.gamma_L / P.atmass;
        if (P.lattlang) {
            const double h = dt / 2.0;
            if (P.localcolmot && n > 0) dp_colmot(lst, n, gm * h / n, false, D);
            const double eb = P.localcolmot ? exp(-gm * h * (1.0 - 1.0 / n)) : exp(-gm * h);
            for (int q = 0; q < n; ++q) {
                int a = lst[q];
                double4 p = D.p[a];
                D.p[a] = make_double4(eb * p.x, eb * p.y, eb * p.z, 0.0);
            }
            const double Tk = P.eltemp ? D.Te[c] : P.temperature;
            const double len = sqrt(2.0 * Tk * P.gamma_L / dt);
            double3 fsum = make_double3(0.0, 0.0, 0.0);
            for (int q = 0; q < n; ++q) {
                int a = lst[q];
                double3 ff = mul3(len, normal3(P.seed, P.replica, a, pass));
                double3 pa = add3(d3(D.p[a]), mul3(dt, add3(d3(D.f[a]), ff)));
                D.p[a] = make_double4(pa.x, pa.y, pa.z, 0.0);
                fsum = add3(fsum, ff);
            }
            if (P.localcolmot && n > 0) {
                double3 ave = make_double3(fsum.x / n, fsum.y / n, fsum.z / n);
                for (int q = 0; q < n; ++q) {
                    int a = lst[q];
                    double3 pa = sub3(d3(D.p[a]), mul3(dt, ave));
                    D.p[a] = make_double4(pa.x, pa.y, pa.z, 0.0);
                }
            }
            for (int q = 0; q < n; ++q) {
                int a = lst[q];
                double4 p = D.p[a];
                D.p[a] = make_double4(eb * p.x, eb * p.y, eb * p.z, 0.0);
            }
            if (P.localcolmot && n > 0) dp_colmot(lst, n, gm * h / n, true, D);
        } else {
            for (int q = 0; q < n; ++q) {
                int a = lst[q];
                double3 pa = add3(d3(D.p[a]), mul3(dt, d3(D.f[a])));
                D.p[a] = make_double4(pa.x, pa.y, pa.z, 0.0);
            }
        }
        for (int q = 0; q < n; ++q) {
            double4 p = D.p[lst[q]];
            acc[0] += p.x; acc[1] += p.y; acc[2] += p.z;
        }
    }
    block_partials<3>(acc, partials);
}

// rescale_momentum, part 1: remove the mean momentum (core_dp_CPU.cpp:270-302)
__global__ __launch_bounds__(256) void k_dp_center(Phys P, Dev D, const double* psum, double* partials) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (a < D.N) {
        const double inv_n = 1.0 / D.N;
        double3 mean = make_double3(psum[0] * inv_n, psum[1] * inv_n, psum[2] * inv_n);
        double3 p = d3(D.p[a]);
        acc[0] = sq3(p) / 2.0 / P.atmass;
        p = sub3(p, mean);
        acc[1] = sq3(p) / 2.0 / P.atmass;
        D.p[a] = make_double4(p.x, p.y, p.z, 0.0);
    }
    block_partials<2>(acc, partials);
}

// rescale_momentum, part 2, and the kinetic energies (core_dp_CPU.cpp:78-84,304-310)
__global__ __launch_bounds__(256) void k_dp_scale(Phys P, Dev D, const double* ke2) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    double factor = sqrt(ke2[0] / ke2[1]);
    if (ke2[1] < 1e-10) factor = 1.0;
    double3 p = mul3(factor, d3(D.p[a]));
    D.p[a] = make_double4(p.x, p.y, p.z, 0.0);
    D.ke[a] = sq3(p) / 2.0 / P.atmass;
}

// core_dr_CPU.cpp:33-44
__global__ __launch_bounds__(256) void k_dr(double dt, Box b, Phys P, Dev D) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    double k = dt / P.atmass;
    double3 r = add3(d3(D.r[a]), mul3(k, d3(D.p[a])));
    r = wrap(b, r);
    D.r[a] = make_double4(r.x, r.y, r.z, 0.0);
}

// ---------------------------------------------------------------------------
// electron subsystem

// calculate_temperature_CPU.cpp:35-148, one thread per link cell
__global__ __launch_bounds__(128) void k_temperature(Tables T, Phys P, Dev D) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.C) return;
    const int n = D.cn[c];
    const int* lst = D.cl + c * kCellCap;
    double sum_ke = 0.0, rup = 0.0, rdn = 0.0, lup = 0.0, ldn = 0.0;
    double3 ap = make_double3(0.0, 0.0, 0.0);
    for (int q = 0; q < n; ++q) {
        int a = lst[q];
        sum_ke += D.ke[a];
        if (P.localcolmot) ap = add3(ap, d3(D.p[a]));
        double3 H = make_double3(0.0, 0.0, 0.0);
        double S = 0.0;
        const int nm = D.nbmn[a];
        const int* nbm = D.nbm + (size_t)a * kMaxMag;
        const double* nbJ = D.nbJ + (size_t)a * kMaxMag;
        for (int k = 0; k < nm; ++k) {
            double J = nbJ[k];
            double4 sj = D.s[nbm[k]];
            H.x += J * sj.x; H.y += J * sj.y; H.z += J * sj.z;
            S += J * sj.w;
        }
        double3 s = d3(D.s[a]);
        rup += sq3(cross3(s, H));
        rdn += dot3(s, H);
        int n0;
        double ratio;
        interp_index(D.rho[a], T.in_over_rhomax, n0, ratio);
        double A = lerp_tab(T.rhotab, kRhoStride, H_A, n0, ratio);
        double B = lerp_tab(T.rhotab, kRhoStride, H_B, n0, ratio);
        double C = lerp_tab(T.rhotab, kRhoStride, H_C, n0, ratio);
        double Dd = lerp_tab(T.rhotab, kRhoStride, H_D, n0, ratio);
        double s2 = sq3(s), s4 = s2 * s2;
        double s0 = sqrt(s2);
        double k = -(2.0 * A + 4.0 * B * s2 + 6.0 * C * s4 + 8.0 * Dd * s4 * s2);
        double hc = -S / s0;
        double3 L = make_double3((k + hc) * s.x, (k + hc) * s.y, (k + hc) * s.z);
        lup += sq3(add3(H, L));
        ldn += 6.0 * A + 20.0 * B * s2 + 42.0 * C * s4 + 72.0 * Dd * s4 * s2;
        ldn += 2.0 * S / s0;
        D.hh[a] = make_double4(H.x, H.y, H.z, 0.0);
        D.hl[a] = make_double4(L.x, L.y, L.z, 0.0);
        D.sjs[a] = S;
        D.s[a].w = s0;
    }
    const double apc = (double)D.N / (double)D.C;
    const double gm = P.gamma_L / P.atmass;
    if (P.localcolmot) {
        ap = make_double3(ap.x / n, ap.y / n, ap.z / n);
        sum_ke -= n * sq3(ap) / 2.0 / P.atmass;
        D.Tl[c] = 2.0 / 3.0 * sum_ke / (n - 1.0) / kBoltz;
        D.Gel[c] = 3.0 * (n - 1.0) * kBoltz * gm / apc;
        D.cellv[c] = make_double4(ap.x, ap.y, ap.z, 0.0);
    } else {
        D.Tl[c] = sum_ke * 2.0 / 3.0 / n / kBoltz;
        D.Gel[c] = 3.0 * n * kBoltz * gm / apc;
    }
    D.TsR[c] = rup / rdn / 2.0 / kBoltz;
    D.TsL[c] = lup / ldn / kBoltz;
    D.Ges[c] = kBoltz * P.gamma_S * ldn / apc;
    D.sRup[c] = rup; D.sRdn[c] = rdn; D.sLup[c] = lup; D.sLdn[c] = ldn;
}

// core_dTe sub-step, first half: A(h), B(h), then Ee = E(Te) as C starts
// (core_dTe_CPU.cpp:51-55,66-129,133-140)
__device__ __forceinline__ void dte_A(int c, double h, const Phys& P, const Dev& D) {
    double ee = te_to_ee(D.Te[c]);
    double dE = 0.0;
    if (P.lattlang) dE += D.Gel[c] * D.Tl[c];
    if (P.spinlang) dE += D.Ges[c] * D.TsL[c];
    ee += h * dE;
    D.Ee[c] = ee;
    D.Te[c] = ee_to_te(ee);
}

__device__ __forceinline__ void dte_B(int c, double h, const Phys& P, const Dev& D) {
    double G = 0.0;
    if (P.lattlang) G += D.Gel[c];
    if (P.spinlang) G += D.Ges[c];
    G *= h / kBoltz;
    double T0 = D.Te[c];
    double k1 = T0 * (exp(-G / ce_of(T0)) - 1.0) / h;
    double T1 = T0 + k1 * h / 2.0;
    double k2 = T1 * (exp(-G / ce_of(T1)) - 1.0) / h;
    double T2 = T0 + k2 * h / 2.0;
    double k3 = T2 * (exp(-G / ce_of(T2)) - 1.0) / h;
    double T3 = T0 + k3 * h;
    double k4 = T3 * (exp(-G / ce_of(T3)) - 1.0) / h;
    D.Te[c] = T0 + h / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
}

__global__ __launch_bounds__(256) void k_dte_ab(double h, Phys P, Dev D, double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (c < D.C) {
        dte_A(c, h, P, D);
        dte_B(c, h, P, D);
        double ee = te_to_ee(D.Te[c]);
        D.Ee[c] = ee;
        acc[0] = ee;
    }
    block_partials<1>(acc, partials);
}

// one dependency level of the C diffusion sweep (core_dTe_CPU.cpp:155-197)
__global__ __launch_bounds__(256) void k_dte_sweep(const int* cells, int ncell, Box b, double pref,
                                                   double dx2, double dy2, double dz2, Dev D) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    int c = cells[t];
    const double* Te = D.Te;
    double tc = Te[c];
    double ee = te_to_ee(tc);
    ee += pref * ((Te[neigh_cell(b, c, 0)] + Te[neigh_cell(b, c, 13)] - 2.0 * tc) / dx2 +
                  (Te[neigh_cell(b, c, 2)] + Te[neigh_cell(b, c, 14)] - 2.0 * tc) / dy2 +
                  (Te[neigh_cell(b, c, 12)] + Te[neigh_cell(b, c, 15)] - 2.0 * tc) / dz2);
    D.Ee[c] = ee;
    D.Te[c] = ee_to_te(ee);
}

__global__ __launch_bounds__(256) void k_dte_end(Dev D, double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (c < D.C) {
        double ee = te_to_ee(D.Te[c]);
        D.Ee[c] = ee;
        acc[0] = ee;
    }
    block_partials<1>(acc, partials);
}

// renormalise to the starting total (core_dTe_CPU.cpp:199-216), then B(h), A(h)
__global__ __launch_bounds__(256) void k_dte_finish(double h, Phys P, Dev D, const double* se) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.C) return;
    double f = se[0] / se[1];
    double ee = D.Ee[c] * f;
    D.Ee[c] = ee;
    D.Te[c] = ee_to_te(ee);
    dte_B(c, h, P, D);
    dte_A(c, h, P, D);
}

// ---------------------------------------------------------------------------
// observables (check_energy_CPU.cpp, check_temperature_CPU.cpp, check_spin_CPU.cpp)

__global__ __launch_bounds__(256) void k_obs_atoms(Tables T, Phys P, Dev D, double* partials,
                                                   int scan_spin) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[kObsAtomK + 1];
#pragma unroll
    for (int k = 0; k <= kObsAtomK; ++k) acc[k] = 0.0;
    if (a < D.N) {
        acc[0] = D.pe[a];
        acc[1] = D.ke[a];
        acc[2] = D.me[a];
        acc[3] = D.me0[a];
        double3 s = d3(D.s[a]);
        acc[4] = s.x; acc[5] = s.y; acc[6] = s.z;
        acc[11] = D.vir[a];
        if (scan_spin) {
            // check_temperature_CPU.cpp:64-77,97-128 (no eltemp)
            double3 H = make_double3(0.0, 0.0, 0.0);
            double S = 0.0;
            const int nm = D.nbmn[a];
            const int* nbm = D.nbm + (size_t)a * kMaxMag;
            const double* nbJ = D.nbJ + (size_t)a * kMaxMag;
            for (int k = 0; k < nm; ++k) {
                double J = nbJ[k];
                double4 sj = D.s[nbm[k]];
                H.x += J * sj.x; H.y += J * sj.y; H.z += J * sj.z;
                S += J * sj.w;
            }
            acc[7] = sq3(cross3(s, H));
            acc[8] = dot3(s, H);
            int n0;
            double ratio;
            interp_index(D.rho[a], T.in_over_rhomax, n0, ratio);
            double A = lerp_tab(T.rhotab, kRhoStride, H_A, n0, ratio);
            double B = lerp_tab(T.rhotab, kRhoStride, H_B, n0, ratio);
            double C = lerp_tab(T.rhotab, kRhoStride, H_C, n0, ratio);
            double Dd = lerp_tab(T.rhotab, kRhoStride, H_D, n0, ratio);
            double s2 = sq3(s), s4 = s2 * s2;
            double s0 = sqrt(s2);
            double k = -(2.0 * A + 4.0 * B * s2 + 6.0 * C * s4 + 8.0 * Dd * s4 * s2);
            double hc = -S / s0;
            double3 L = make_double3((k + hc) * s.x, (k + hc) * s.y, (k + hc) * s.z);
            acc[9] = sq3(add3(H, L));
            acc[10] = 2.0 * S / s0 + (6.0 * A + 20.0 * B * s2 + 42.0 * C * s4 + 72.0 * Dd * s4 * s2);
            D.hh[a] = make_double4(H.x, H.y, H.z, 0.0);
            D.hl[a] = make_double4(L.x, L.y, L.z, 0.0);
            D.sjs[a] = S;
        }
    }
    block_partials<kObsAtomK + 1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_obs_cells(Dev D, double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[kObsCellK] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (c < D.C) {
        double ee = te_to_ee(D.Te[c]);
        D.Ee[c] = ee;
        acc[0] = ee;
        acc[1] = D.sRup[c]; acc[2] = D.sRdn[c]; acc[3] = D.sLup[c]; acc[4] = D.sLdn[c];
        acc[5] = D.Te[c];
    }
    block_partials<kObsCellK>(acc, partials);
}

// renormalizeEnergy feedback (check_energy_CPU.cpp:103-110); the printed mean
// Te is taken afterwards (check_temperature_CPU.cpp:130-134), so the new
// temperatures are summed here.
__global__ __launch_bounds__(256) void k_renorm(Dev D, const double* err, double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (c < D.C) {
        double ee = D.Ee[c] - err[0];
        D.Ee[c] = ee;
        double te = ee_to_te(ee);
        D.Te[c] = te;
        acc[0] = te;
    }
    block_partials<1>(acc, partials);
}

__global__ __launch_bounds__(256) void k_check_nonfinite(Dev D) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    double4 r = D.r[a], p = D.p[a], s = D.s[a];
    double x = r.x + r.y + r.z + p.x + p.y + p.z + s.x + s.y + s.z;
    if (!isfinite(x)) atomicOr(D.err, kErrNonFinite);
}

}  // namespace sld
