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

// sum over the 64 lanes of a wave, by a fixed butterfly of shuffles
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}


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
        for (int k = 0; k < NK; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = sh[k][0];
    }
}

// out[k] = sum over blocks of partials[k*nblk + b]; one block of 256 threads.
__global__ __launch_bounds__(256) void k_reduce(const double* partials, int nblk, int nk, double* out) {
    __shared__ double sh[256];
    const int t = threadIdx.x;
    for (int k = 0; k < nk; ++k) {
        double acc = 0.0;
        for (int b = t; b < nblk; b += 256) acc += partials[(size_t)k * nblk + b];
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

// Per-replica sums: block r of the grid (1,024 threads) reduces partial rows
// [seg[r], seg[r+1]) into out[r * kScal + out_off + k], in an order fixed by
// the row count alone: thread t adds rows t, t + 1024, ... in turn, each wave
// sums by shuffles, thread 0 adds the 16 wave sums in order.
__global__ __launch_bounds__(1024) void k_reduce_seg(const double* partials, const int* seg, int nk, double* out,
                                                     int out_off) {
    constexpr int KG = 4;  // components summed together
    __shared__ double sh[KG][16];
    const int t = threadIdx.x;
    const int b0 = seg[blockIdx.x], nb = seg[blockIdx.x + 1] - b0;
    const size_t ntot = (size_t)seg[gridDim.x];  // rows of each component
    for (int k0 = 0; k0 < nk; k0 += KG) {
        const int kk = min(KG, nk - k0);
        // thread t adds rows t + 1024 u in order, eight rows of up to four
        // components in flight
        double acc[KG] = {0.0, 0.0, 0.0, 0.0};
        for (int c0 = t; c0 < nb; c0 += 8 * 1024) {
            double v[8][KG];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int b = c0 + u * 1024;
#pragma unroll
                for (int q = 0; q < KG; ++q)
                    v[u][q] = b < nb && q < kk ? partials[(k0 + q) * ntot + b0 + b] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int q = 0; q < KG; ++q) acc[q] += v[u][q];
        }
#pragma unroll
        for (int q = 0; q < KG; ++q) {
            acc[q] = wave_sum(acc[q]);
            if ((t & 63) == 0) sh[q][t >> 6] = acc[q];
        }
        __syncthreads();
        if (t < kk) {
            double a = 0.0;
            for (int w = 0; w < 16; ++w) a += sh[t][w];
            out[(size_t)blockIdx.x * kScal + out_off + k0 + t] = a;
        }
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
__global__ __launch_bounds__(256) void k_bin(Dev D) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    const int rr = D.arep[a];
    if (rr < 0) return;
    const Rep& R = D.rep[rr];
    const int old = D.cell[a];
    const int lc = bin_cell(R.box, d3(D.r[a]));
    const int c = R.cell0 + lc;
    D.cell_prev[a] = old;
    D.cell[a] = c;
    if (c != old) {
        if (!cells_adjacent(R.box, old - R.cell0, lc)) atomicOr(D.err, kErrFarMove);
        const int slot = atomicAdd(D.inc_n + c, 1);
        if (slot < kCellCap) D.inc[c * kCellCap + slot] = a;
    }
}

// Rebuild cell c's list. The reference detaches every mover and pushes it at
// the head of its new cell in ascending atom index (links_CPU.cpp:144-173),
// so a cell's new list is [movers in, descending index] + [stayers, old order].
// k_bin has pushed this step's movers into the cell's incoming list in
// arbitrary order; sorting them by index makes the list deterministic.
__global__ __launch_bounds__(128) void k_relink(const int* cl_old, const int* cn_old, Dev D) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.C) return;
    const int rr = D.crep[c];
    if (rr < 0) return;
    int mv[kCellCap];
    const int ninc = D.inc_n[c];
    bool over = ninc > kCellCap;
    const int nm = over ? kCellCap : ninc;
    for (int k = 0; k < nm; ++k) {
        const int a = D.inc[c * kCellCap + k];
        // insertion keeping descending atom index
        int q = k;
        while (q > 0 && mv[q - 1] < a) { mv[q] = mv[q - 1]; --q; }
        mv[q] = a;
    }
    if (ninc != 0) D.inc_n[c] = 0;
    int* out = D.cl + c * kCellCap;
    int k = 0;
    for (int q = 0; q < nm; ++q) {
        D.cslot[mv[q]] = c * kCellCap + k;
        out[k++] = mv[q];
    }
    int n = cn_old[c];
    for (int q = 0; q < n; ++q) {
        int a = cl_old[c * kCellCap + q];
        if (D.cell[a] == c) {
            if (k < kCellCap) {
                D.cslot[a] = c * kCellCap + k;
                out[k++] = a;
            } else {
                over = true;
            }
        }
    }
    D.cn[c] = k;
    if (over) atomicOr(D.err, kErrCellOverflow);
}

// ---------------------------------------------------------------------------
// neighbour list + density (calculate_rho_CPU.cpp:33-92 for rho, the 27-cell
// scan order of calculate_spin_CPU.cpp:201-235 for the lists)
//
// One wavefront per link cell. The ids of the 27 cells' atoms (slots 0..25,
// then the cell itself, each in list order) are staged in LDS; every lane
// holds up to kCandReg candidate positions in registers and the wave walks
// the cell's atoms, testing all candidates at once. In-range pairs are
// compacted with a ballot and a lane prefix count, so each list keeps the
// reference's scan order.

constexpr int kNbWaves = 4;                 // most waves (cells) per block
constexpr int kMaxCand = 27 * kCellCap;     // upper bound of candidates
constexpr int kCandReg = 4;                 // passes of 64 candidates held in registers


static_assert(kCellCap <= 64, "k_nbuild holds a cell's own atoms one per lane");

// lane l's value of v, as a wave-uniform double
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__global__ __launch_bounds__(256) void k_nbuild(Tables T, Phys P, Dev D, int write_rho) {
    __shared__ int s_cand[kNbWaves][kMaxCand];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int lb = xcd_block((D.C + wpb - 1) / wpb);
    if (lb < 0) return;  // whole padding block, before any barrier
    const int c = lb * wpb + w;
    const int rr = __builtin_amdgcn_readfirstlane(c < D.C ? D.crep[c] : -1);
    const bool valid = rr >= 0;
    Box b{};
    int cell0 = 0;
    if (valid) {
        b = D.rep[rr].box;
        cell0 = D.rep[rr].cell0;
    }
    int cid = 0, cnt = 0;
    if (valid && lane < 27) {
        cid = lane < 26 ? cell0 + neigh_cell(b, c - cell0, lane) : c;
        cnt = D.cn[cid];
    }
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    const int total = __shfl(incl, 26, 64);
    if (valid && lane < 27) {
        const int* src = D.cl + cid * kCellCap;
        const int o = incl - cnt;
        for (int k = 0; k < cnt; ++k) s_cand[w][o + k] = src[k];
    }
    __syncthreads();
    if (!valid) return;
    const int npass = (total + 63) >> 6;
    int jc[kCandReg];
    double4 rc[kCandReg];
#pragma unroll
    for (int q = 0; q < kCandReg; ++q) {
        int idx = q * 64 + lane;
        jc[q] = (q < npass && idx < total) ? s_cand[w][idx] : -1;
        rc[q] = jc[q] >= 0 ? D.r[jc[q]] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int nc = D.cn[c];
    const int* own = D.cl + c * kCellCap;
    bool over = false;
    // the cell's own atoms (nc <= kCellCap <= 64): id and position one per
    // lane, loaded once, then read at the loop's uniform index, so no atom
    // waits for its own id and position before its distances
    const int oi = lane < nc ? own[lane] : 0;
    const double4 orv = lane < nc ? D.r[oi] : make_double4(0.0, 0.0, 0.0, 0.0);
    for (int k = 0; k < nc; ++k) {
        const int i = __builtin_amdgcn_readlane(oi, k);
        const double3 ri = make_double3(readlane_f64(orv.x, k), readlane_f64(orv.y, k), readlane_f64(orv.z, k));
        int nn = 0, nm = 0;
        double rho = 0.0;
        int* nb = D.nb + (size_t)i * kMaxNb;
        int* nbm = D.nbm + (size_t)i * kMaxMag;
        double* nbJ = D.nbJ + (size_t)i * kMaxMag;
        // the passes held in registers: every distance first, then every
        // table row of the pairs within rcut_max requested at once, then the
        // list entries and density terms in pass order
        double rsqv[kCandReg];
        bool inmaxv[kCandReg];
#pragma unroll
        for (int q = 0; q < kCandReg; ++q) {
            const bool ok = q < npass && jc[q] >= 0 && jc[q] != i;
            double rsq = 1e300;
            if (ok) {
                (void)min_image_sq(b, sub3(ri, d3(rc[q])), rsq);
            }
            rsqv[q] = rsq;
            inmaxv[q] = ok && rsq < P.rcut_max_sq;
        }
        double2 sfv[kCandReg], jv[kCandReg];
        double ratv[kCandReg], rv[kCandReg];
#pragma unroll
        for (int q = 0; q < kCandReg; ++q) {
            sfv[q] = make_double2(0.0, 0.0);
            jv[q] = make_double2(0.0, 0.0);
            ratv[q] = 0.0;
            rv[q] = 0.0;
            if (inmaxv[q]) {
                const double r = sqrt(rsqv[q]);
                int n0;
                double ratio;
                interp_index(r, T.in_over_rmax, n0, ratio);
                const double* row = T.prow + (size_t)n0 * kPRowStride;
                sfv[q] = *reinterpret_cast<const double2*>(row + P_SF);
                jv[q] = *reinterpret_cast<const double2*>(row + P_J);
                ratv[q] = ratio;
                rv[q] = r;
            }
        }
#pragma unroll
        for (int q = 0; q < kCandReg; ++q) {
            if (q >= npass) break;
            const int j = jc[q];
            const bool inmax = inmaxv[q];
            const unsigned long long m1 = __ballot(inmax);
            const int pos = nn + __popcll(m1 & lt);
            if (inmax && pos < kMaxNb) nb[pos] = j;
            nn += __popcll(m1);
            const bool inmag = inmax && rsqv[q] < P.rcut_mag_sq;
            const unsigned long long m2 = __ballot(inmag);
            const int pos2 = nm + __popcll(m2 & lt);
            if (inmax) {
                const double ratio = ratv[q];
                if (rv[q] < P.rcut_pot) rho += sfv[q].x * ratio + sfv[q].y * (1.0 - ratio);
                if (inmag && pos2 < kMaxMag) {
                    nbm[pos2] = j;
                    nbJ[pos2] = jv[q].x * ratio + jv[q].y * (1.0 - ratio);
                }
            }
            nm += __popcll(m2);
        }
        // candidates past the registers (more than 64 x kCandReg): one pass at a time
        for (int q = kCandReg; q < npass; ++q) {
            const int idx = q * 64 + lane;
            const int j = idx < total ? s_cand[w][idx] : -1;
            const double4 rj = j >= 0 ? D.r[j] : make_double4(0.0, 0.0, 0.0, 0.0);
            bool ok = j >= 0 && j != i;
            double rsq = 1e300;
            if (ok) {
                (void)min_image_sq(b, sub3(ri, d3(rj)), rsq);
            }
            const bool inmax = ok && rsq < P.rcut_max_sq;
            const unsigned long long m1 = __ballot(inmax);
            const int pos = nn + __popcll(m1 & lt);
            if (inmax && pos < kMaxNb) nb[pos] = j;
            nn += __popcll(m1);
            const bool inmag = inmax && rsq < P.rcut_mag_sq;
            const unsigned long long m2 = __ballot(inmag);
            const int pos2 = nm + __popcll(m2 & lt);
            if (inmax) {
                double r = sqrt(rsq);
                int n0;
                double ratio;
                interp_index(r, T.in_over_rmax, n0, ratio);
                const double* row = T.prow + (size_t)n0 * kPRowStride;
                if (r < P.rcut_pot) rho += row[P_SF] * ratio + row[P_SF + 1] * (1.0 - ratio);
                if (inmag && pos2 < kMaxMag) {
                    nbm[pos2] = j;
                    nbJ[pos2] = row[P_J] * ratio + row[P_J + 1] * (1.0 - ratio);
                }
            }
            nm += __popcll(m2);
        }
        rho = wave_sum(rho);
        if (lane == 0) {
            D.nbn[i] = nn < kMaxNb ? nn : kMaxNb;
            D.nbmn[i] = nm < kMaxMag ? nm : kMaxMag;
            if (write_rho) D.rho[i] = rho;
        }
        over |= nn > kMaxNb || nm > kMaxMag;
    }
    if (over && lane == 0) atomicOr(D.err, kErrNbOverflow);
}

// ---------------------------------------------------------------------------
// force and energy (calculate_force_energy_CPU.cpp:33-301), full neighbour
// list: each atom sums its own pair terms; no atomics.

// dF/drho plus the Landau derivative terms of one atom
// (calculate_force_energy_CPU.cpp:98-107)
__global__ __launch_bounds__(256) void k_fp(Tables T, Dev D) {
    const int lb = xcd_block((D.N + blockDim.x - 1) / blockDim.x);
    if (lb < 0) return;
    int i = lb * blockDim.x + threadIdx.x;
    if (i >= D.N || D.arep[i] < 0) return;
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

constexpr int kForceBatch = 8;  // neighbour pairs whose loads the force pass issues together

__global__ __launch_bounds__(256) void k_force(Tables T, Phys P, Dev D) {
    const int lb = xcd_block((D.N + blockDim.x - 1) / blockDim.x);
    if (lb < 0) return;
    int i = lb * blockDim.x + threadIdx.x;
    if (i >= D.N) return;
    const int rr = D.arep[i];
    if (rr < 0) return;
    const Rep& R = D.rep[rr];
    const Box b = R.box;
    const double3 ri = d3(D.r[i]);
    const double4 si4 = D.s[i];
    const double3 si = d3(si4);
    const double si_sq = sq3(si);
    const double fpi = D.fp[i];
    double3 f = make_double3(0.0, 0.0, 0.0);
    double pe = 0.0, me = 0.0, me0 = 0.0, vir = 0.0;
    double st0 = 0.0, st1 = 0.0, st2 = 0.0, st3 = 0.0, st4 = 0.0, st5 = 0.0;
    double m1 = 0.0, m2 = 0.0;  // localvol: sums of 1/r and 1/r^2 within rcut_vol
    const int nn = D.nbn[i];
    const int* nb = D.nb + (size_t)i * kMaxNb;
    // kForceBatch pairs at a time: their neighbour ids, then their
    // positions, F'(rho) and spins, then their table rows, all in flight
    // together; the pair terms are then added in list order, as a
    // pair-at-a-time loop adds them
    for (int k0 = 0; k0 < nn; k0 += kForceBatch) {
        int jj[kForceBatch];
#pragma unroll
        for (int u = 0; u < kForceBatch; ++u)
            jj[u] = k0 + u < nn ? nb[k0 + u] : i;  // past the list: a valid slot, never added
        double4 rj[kForceBatch], sj4[kForceBatch];
        double fpj[kForceBatch];
#pragma unroll
        for (int u = 0; u < kForceBatch; ++u) {
            rj[u] = D.r[jj[u]];
            fpj[u] = D.fp[jj[u]];
            sj4[u] = D.s[jj[u]];
        }
        double3 rij[kForceBatch];
        double rsq[kForceBatch], r[kForceBatch], ratio[kForceBatch];
        const double* q[kForceBatch];
#pragma unroll
        for (int u = 0; u < kForceBatch; ++u) {
            rij[u] = min_image_sq(b, sub3(ri, d3(rj[u])), rsq[u]);
            r[u] = sqrt(rsq[u]);
            int n0;
            interp_index(r[u], T.in_over_rmax, n0, ratio[u]);
            q[u] = T.prow + (size_t)n0 * kPRowStride;
        }
#pragma unroll
        for (int u = 0; u < kForceBatch; ++u) {
            if (k0 + u >= nn) break;
            const double* qq = q[u];
            const double rt = ratio[u], wr = 1.0 - rt;
            double dudr = 0.0;
            if (r[u] < P.rcut_pot) {
                double dsf = qq[P_DSF] * rt + qq[P_DSF + 1] * wr;
                double dpr = qq[P_DPR] * rt + qq[P_DPR + 1] * wr;
                double pr = qq[P_PR] * rt + qq[P_PR + 1] * wr;
                dudr = (fpi + fpj[u]) * dsf + dpr;
                pe += 0.5 * pr;
            }
            if (r[u] < P.rcut_mag) {
                double3 sj = d3(sj4[u]);
                double sdot = dot3(si, sj);
                double stimes = sqrt(si_sq * sq3(sj));
                double J = qq[P_J] * rt + qq[P_J + 1] * wr;
                double dJ = qq[P_DJ] * rt + qq[P_DJ + 1] * wr;
                dudr += -dJ * (sdot - stimes);
                double Jh = J / 2.0;
                me0 += Jh * stimes;
                me += -Jh * sdot;
            }
            double force = -dudr / r[u];
            double3 fij = mul3(force, rij[u]);
            f = add3(f, fij);
            st0 += fij.x * rij[u].x; st1 += fij.y * rij[u].y; st2 += fij.z * rij[u].z;
            st3 += fij.x * rij[u].y; st4 += fij.y * rij[u].z; st5 += fij.z * rij[u].x;
            vir += -force * rsq[u];
            if (P.localvol && r[u] < P.rcut_vol) {  // calculate_force_energy_CPU.cpp:80-90
                m1 += 1.0 / r[u];
                m2 += 1.0 / rsq[u];
            }
        }
    }
    // one-body terms (calculate_force_energy_CPU.cpp:196-215)
    int n0;
    double ratio;
    interp_index(D.rho[i], T.in_over_rhomax, n0, ratio);
    double F = lerp_tab(T.rhotab, kRhoStride, H_F, n0, ratio);
    const double4 lan = D.lan[i];
    const double A = lan.x, B = lan.y, C = lan.z, Dd = lan.w;
    double s4 = si_sq * si_sq;
    pe += F;
    if (P.extfield) me -= dot3(si, d3(D.hext[i]));  // calculate_force_energy_CPU.cpp:199-201
    me += A * si_sq + B * s4 + C * s4 * si_sq + Dd * s4 * s4;
    D.f[i] = make_double4(f.x, f.y, f.z, 0.0);
    D.pe[i] = pe;
    D.me[i] = me;
    D.me0[i] = me0;
    // the reference books -force*r^2 once per pair on the half-shell owner;
    // here each atom keeps half of each of its pairs (same total virial)
    D.vir[i] = 0.5 * vir;
    // per-atom stress with the kinetic part (calculate_force_energy_CPU.cpp:283-293);
    // with localvol the local volume (:270-276) is normalised over the
    // replica first, so the numerators are stored and k_lvol_apply divides
    const double m = P.atmass;
    const double4 p = D.p[i];
    const size_t N = D.N;
    double num[6] = {p.x * p.x / m + st0 / 2.0,   p.y * p.y / m + st1 / 2.0,   p.z * p.z / m + st2 / 2.0,
                     (p.x * p.y) / m + st3 / 2.0, (p.y * p.z) / m + st4 / 2.0, (p.z * p.x) / m + st5 / 2.0};
    if (P.localvol) {
        const double local_radius = 0.5 * m1 / m2;
        D.lvol[i] = 4.0 * kPi / 3.0 * pow(local_radius, 3.0);
#pragma unroll
        for (int k = 0; k < 6; ++k) D.stress[k * N + i] = num[k];
    } else {
        const double lv = R.box_volume / R.natom;
#pragma unroll
        for (int k = 0; k < 6; ++k) D.stress[k * N + i] = num[k] / lv;
    }
}

// localvol: each replica's sum of the local volumes, one partial per 256
// atom slots (calculate_force_energy_CPU.cpp:270-275)
__global__ __launch_bounds__(256) void k_lvol_sum(Dev D, double* partials) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (a < D.N && D.arep[a] >= 0) acc[0] = D.lvol[a];
    block_partials<1>(acc, partials);
}

// localvol: the volumes scaled to fill the box, then each atom's stress over
// its volume (calculate_force_energy_CPU.cpp:279-293)
__global__ __launch_bounds__(256) void k_lvol_apply(Dev D, const double* scal) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    const int rr = D.arep[a];
    if (rr < 0) return;
    const double volume_factor = D.rep[rr].box_volume / scal[(size_t)rr * kScal + kScalLvol];
    const double lv = D.lvol[a] * volume_factor;
    D.lvol[a] = lv;
    const size_t N = D.N;
#pragma unroll
    for (int k = 0; k < 6; ++k) D.stress[k * N + a] = D.stress[k * N + a] / lv;
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
    // -sjs / |s| through one reciprocal square root (the reference's sqrt and
    // division, to about an ulp, on a shorter dependent chain)
    double hc = s2 > 0.0 ? -sjs * rsqrt(s2) : 0.0;
    return make_double3((k + hc) * s.x, (k + hc) * s.y, (k + hc) * s.z);
}

// Exchange field H = Hext + sum J s_j (Hext with extfield, else 0:
// calculate_spin_CPU.cpp:35-39) and S = sum J |s_j| over the rcut_mag list,
// gathered eight neighbours at a time with all loads of a batch in flight.
__device__ __forceinline__ void exchange_field(int a, const Dev& D, const Phys& P, double3& H, double& S) {
    const int nm = D.nbmn[a];
    const int* ids = D.nbm + (size_t)a * kMaxMag;
    const double* js = D.nbJ + (size_t)a * kMaxMag;
    double hx = 0.0, hy = 0.0, hz = 0.0, ss = 0.0;
    if (P.extfield) {
        const double4 he = D.hext[a];
        hx = he.x; hy = he.y; hz = he.z;
    }
#pragma unroll 1
    for (int b0 = 0; b0 < nm; b0 += 8) {
        const int4 i0 = *reinterpret_cast<const int4*>(ids + b0);
        const int4 i1 = *reinterpret_cast<const int4*>(ids + b0 + 4);
        const int id[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
        double J[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double2 v = *reinterpret_cast<const double2*>(js + b0 + 2 * q);
            J[2 * q] = v.x; J[2 * q + 1] = v.y;
        }
        double4 sj[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) sj[q] = D.s[b0 + q < nm ? id[q] : a];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double Jq = b0 + q < nm ? J[q] : 0.0;
            hx += Jq * sj[q].x; hy += Jq * sj[q].y; hz += Jq * sj[q].z;
            ss += Jq * sj[q].w;
        }
    }
    H = make_double3(hx, hy, hz);
    S = ss;
}

// One dependency level of a core_ds sweep: each listed cell walks its list
// head to tail (forward) or tail to head (backward)
// (calculate_spin_CPU.cpp:33-235, core_ds_CPU.cpp:35-83).
//
// A level holds a few thousand to forty thousand cells, and each cell's walk
// is a chain of dependent work, so the level lasts as long as one chain.
// Everything that does not depend on other atoms' new spins is taken off that
// chain:
// - the Langevin normals of the pass come from k_spin_noise, a fully
//   parallel kernel launched before the sweep;
// - the per-atom work is software pipelined: during the update of atom q the
//   lanes already hold atom q+1's neighbour ids and J values, have issued
//   the loads of those neighbours' spins and of atom q+1's own spin, noise
//   and Landau coefficients, and have atom q+2's ids and density in flight.
// The only spin such a load can miss is atom q's own new value (written after
// the load was issued). The field is summed over the loaded spins and then
// corrected by J_q (s_q,new - s_q,old), where J_q is atom q's weight in the
// list (zero if it is not a neighbour), so every atom sees the spins the
// sequential sweep shows it, up to the rounding of that correction.
//
// The field sums H = sum J_e s_e and S = sum J_e |s_e| run over the
// atom's rcut_mag list as four partial sums: P_r adds the entries 4r..4r+3,
// then 16+4r..16+4r+3, in order, and H = Hext + ((P0 + P1) + (P2 + P3)).
// A cell's walk is spread over LPC = 1, 2 or 4 lanes: lane j of the cell
// holds the partials r = j, j + LPC, ..., gathers their entries, and the
// lanes combine them by shuffles in that same order; every lane then runs the
// per-atom update (the same arithmetic on the same values, so the same bits)
// and lane 0 stores the result. More lanes per cell put more waves on a level
// (a 39,304-cell level fills 614 waves with one lane per cell, 2,456 with four)
// and keep more neighbour loads in flight per atom; the results are bitwise
// the same for every LPC.

constexpr int kPipe = 16;      // rcut_mag neighbours loaded ahead (more: loaded when used)
constexpr int kMagPart = 4;    // partial sums of the field, 4 entries per chunk

// rotation of s about Heff by |Heff| dt / hbar (core_ds_CPU.cpp:90-127); the
// axis and angle are computed once for both half-rotations of a step
struct Rot {
    double wx, wy, wz, A, B;
};

__device__ __forceinline__ Rot make_rot(double3 H, double dt) {
    constexpr double m_inv_hbar = -1.0 / kHbar;
    double3 w = make_double3(H.x * m_inv_hbar, H.y * m_inv_hbar, H.z * m_inv_hbar);
    double w0 = sqrt(sq3(w));
    Rot r;
    if (w0 > 0.0) {
        const double iw = 1.0 / w0;
        r.wx = w.x * iw; r.wy = w.y * iw; r.wz = w.z * iw;
    } else {
        r.wx = 0.0; r.wy = 0.0; r.wz = 0.0;
    }
    double sn, cs;
    sincos(w0 * dt, &sn, &cs);
    r.A = sn;
    r.B = 1.0 - cs;
    return r;
}

__device__ __forceinline__ double3 apply_rot(const Rot& r, double3 s) {
    double o12 = r.wx * r.wy, o23 = r.wy * r.wz, o13 = r.wx * r.wz;
    double o1s = r.wx * r.wx, o2s = r.wy * r.wy, o3s = r.wz * r.wz;
    double A = r.A, B = r.B;
    double3 t;
    t.x = s.x + (s.x * B * (-o2s - o3s) + s.y * (B * o12 - A * r.wz) + s.z * (A * r.wy + B * o13));
    t.y = s.y + (s.y * B * (-o1s - o3s) + s.z * (B * o23 - A * r.wx) + s.x * (A * r.wz + B * o12));
    t.z = s.z + (s.z * B * (-o1s - o2s) + s.x * (B * o13 - A * r.wy) + s.y * (A * r.wx + B * o23));
    return t;
}

// The Langevin normals of one spin pass for every atom (the counter-based
// stream of normal3, keyed by replica seed and stream id, counted by the
// atom's index in its replica and the pass). The pass number is the device
// counter plus an offset fixed by the pass's place in the step, so that the
// launches of a step are the same in every step.
__global__ __launch_bounds__(256) void k_spin_noise(Dev D, int pass_off) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.N) return;
    const int rr = D.arep[i];
    if (rr < 0) return;
    const Rep& R = D.rep[rr];
    const uint64_t pass = D.pass_base[0] + (uint64_t)pass_off;
    const double3 z = normal3(R.seed, R.replica, i - R.atom0, pass);
    D.zn[i] = make_double4(z.x, z.y, z.z, 0.0);
}

// advances the device noise-pass counter by the passes a step consumed
__global__ void k_pass_advance(Dev D, int n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) D.pass_base[0] += (uint64_t)n;
}

// Landau coefficients A..D of an atom: interpolation weights of its density,
// then the two table rows (calculate_spin_CPU.cpp:118-121)
struct LanRows {
    double4 lo, hi;
    double ratio;
};

__device__ __forceinline__ LanRows landau_rows(const Tables& T, double rho) {
    int n0;
    double ratio;
    interp_index(rho, T.in_over_rhomax, n0, ratio);
    // columns A..D sit 16-byte aligned in 128-byte rows
    const double2* lo = reinterpret_cast<const double2*>(T.rhotab + (size_t)(n0 - 1) * kRhoStride + H_A);
    const double2* hi = reinterpret_cast<const double2*>(T.rhotab + (size_t)n0 * kRhoStride + H_A);
    const double2 l0 = lo[0], l1 = lo[1], h0 = hi[0], h1 = hi[1];
    LanRows L;
    L.lo = make_double4(l0.x, l0.y, l1.x, l1.y);
    L.hi = make_double4(h0.x, h0.y, h1.x, h1.y);
    L.ratio = ratio;
    return L;
}

// The Landau coefficients A..D of every atom at its current density
// (calculate_spin_CPU.cpp:118-121), interpolated as the reference does. The
// density changes once a step (calculate_rho); the four spin sweeps, the
// force pass and the temperature and energy observables in between read the
// coefficients from here, 32 contiguous bytes per atom, instead of two
// 128-byte table rows each.
__global__ __launch_bounds__(256) void k_landau(Tables T, Dev D) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.N || D.arep[i] < 0) return;
    const LanRows L = landau_rows(T, D.rho[i]);
    const double w = L.ratio, v = 1.0 - L.ratio;
    D.lan[i] = make_double4(L.lo.x * w + L.hi.x * v, L.lo.y * w + L.hi.y * v, L.lo.z * w + L.hi.z * v,
                            L.lo.w * w + L.hi.w * v);
}

template <int LPC>
__global__ __launch_bounds__(256) void k_spin_phase(const int* cells, int ncell, int backward, int store_fields,
                                                    double dt, Tables T, Phys P, Dev D) {
    constexpr int NCH = kMagPart / LPC;    // chunks of this lane: r = j + LPC * ch
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = gt / LPC;
    const int j = gt & (LPC - 1);
    if (g >= ncell) return;
    const int c = cells[g];
    const int n = D.cn[c];
    if (n == 0) return;
    const int* lst = D.cl + c * kCellCap;
    const bool lang = P.spinlang != 0;
    const double Tk = lang ? (P.eltemp ? D.Te[c] : D.rep[D.crep[c]].temperature) : 0.0;
    const double dth = dt / 2.0, dtq = dt / 4.0, gS = P.gamma_S;
    const double rs = lang ? sqrt(2.0 * Tk * gS / dt) : 0.0;
    auto atom_at = [&](int q) { return lst[backward ? n - 1 - q : q]; };
    auto load_ids = [&](int a, int (&id)[NCH][4]) {
        const int4* p = reinterpret_cast<const int4*>(D.nbm + (size_t)a * kMaxMag);
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const int4 v = p[j + LPC * ch];
            id[ch][0] = v.x; id[ch][1] = v.y; id[ch][2] = v.z; id[ch][3] = v.w;
        }
    };
    auto load_js = [&](int a, double (&J)[NCH][4]) {
        const double2* p = reinterpret_cast<const double2*>(D.nbJ + (size_t)a * kMaxMag);
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const double2 u = p[2 * (j + LPC * ch)], v = p[2 * (j + LPC * ch) + 1];
            J[ch][0] = u.x; J[ch][1] = u.y; J[ch][2] = v.x; J[ch][3] = v.y;
        }
    };
    auto gather = [&](int a, int nm, const int (&id)[NCH][4], double4 (&sj)[NCH][4]) {
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
            for (int e = 0; e < 4; ++e) sj[ch][e] = D.s[4 * (j + LPC * ch) + e < nm ? id[ch][e] : a];
    };
    // sum over the lanes of one cell, in the order (P0 + P1) + (P2 + P3)
    auto combine = [&](const double (&v)[NCH]) {
        if constexpr (LPC == 1) return (v[0] + v[1]) + (v[2] + v[3]);
        if constexpr (LPC == 2) {
            const double u = v[0] + __shfl_xor(v[0], 1);
            const double w = v[1] + __shfl_xor(v[1], 1);
            return u + w;
        }
        const double u = v[0] + __shfl_xor(v[0], 1);
        return u + __shfl_xor(u, 2);
    };

    // pipeline registers: atom q (_c), atom q+1 (_n) and atom q+2's list entry
    int a_c = atom_at(0), nm_c = D.nbmn[a_c];
    int id_c[NCH][4], id_n[NCH][4];
    double J_c[NCH][4];
    double4 sj[NCH][4];
    load_ids(a_c, id_c);
    load_js(a_c, J_c);
    gather(a_c, nm_c, id_c, sj);
    const bool ext = P.extfield != 0;
    double4 own_c = D.s[a_c];
    double4 he_c = ext ? D.hext[a_c] : make_double4(0.0, 0.0, 0.0, 0.0);
    double4 z_c = lang ? D.zn[a_c] : make_double4(0.0, 0.0, 0.0, 0.0);
    double4 lan_c = lang ? D.lan[a_c] : make_double4(0.0, 0.0, 0.0, 0.0);
    int a_n = a_c, nm_n = nm_c;
    if (n > 1) {
        a_n = atom_at(1);
        nm_n = D.nbmn[a_n];
        load_ids(a_n, id_n);
    }
    // the list entry two atoms ahead, so that no iteration waits for one
    // before it can request the next atom's ids
    int a_nn = n > 2 ? atom_at(2) : a_c;
    int prev_a = -1;
    double4 prev_s = make_double4(0.0, 0.0, 0.0, 0.0);
    double4 prev_old = prev_s;  // the previous atom's spin before its update
    for (int q = 0; q < n; ++q) {
        const int a = a_c;
        // this lane's partial sums over the gathered spins (then its
        // overflow entries, loaded now), and the previous atom's weight
        double px[NCH], py[NCH], pz[NCH], ps[NCH];
        double Jp = 0.0;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const int k0 = 4 * (j + LPC * ch);
            double x = 0.0, y = 0.0, z = 0.0, w = 0.0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double4 v = sj[ch][e];
                const double Jk = k0 + e < nm_c ? J_c[ch][e] : 0.0;
                x += Jk * v.x; y += Jk * v.y; z += Jk * v.z;
                w += Jk * v.w;
                Jp = (k0 + e < nm_c && id_c[ch][e] == prev_a) ? Jk : Jp;
            }
            for (int k = kPipe + k0; k < kPipe + k0 + 4 && k < nm_c; ++k) {
                const double Jk = D.nbJ[(size_t)a * kMaxMag + k];
                const double4 v = D.s[D.nbm[(size_t)a * kMaxMag + k]];
                x += Jk * v.x; y += Jk * v.y; z += Jk * v.z;
                w += Jk * v.w;
            }
            px[ch] = x; py[ch] = y; pz[ch] = z; ps[ch] = w;
        }
        if constexpr (LPC >= 2) Jp += __shfl_xor(Jp, 1);
        if constexpr (LPC >= 4) Jp += __shfl_xor(Jp, 2);
        double hx = he_c.x + combine(px), hy = he_c.y + combine(py), hz = he_c.z + combine(pz);
        double S = combine(ps);
        hx += Jp * (prev_s.x - prev_old.x);
        hy += Jp * (prev_s.y - prev_old.y);
        hz += Jp * (prev_s.z - prev_old.z);
        S += Jp * (prev_s.w - prev_old.w);
        const double3 H = make_double3(hx, hy, hz);
        prev_old = own_c;
        double3 s = d3(own_c);
        const double3 zq = d3(z_c);
        const double4 lan = lan_c;
        // pipeline: issue atom q+1's neighbour spins, J values, own spin,
        // noise and Landau rows, and atom q+2's ids and density
        if (q + 1 < n) {
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
                for (int e = 0; e < 4; ++e) id_c[ch][e] = id_n[ch][e];
            gather(a_n, nm_n, id_n, sj);
            load_js(a_n, J_c);
            own_c = D.s[a_n];
            if (ext) he_c = D.hext[a_n];
            if (lang) {
                z_c = D.zn[a_n];
                lan_c = D.lan[a_n];
            }
            a_c = a_n;
            nm_c = nm_n;
            if (q + 2 < n) {
                a_n = a_nn;
                nm_n = D.nbmn[a_n];
                load_ids(a_n, id_n);
                if (q + 3 < n) a_nn = atom_at(q + 3);
            }
        }
        double3 L = make_double3(0.0, 0.0, 0.0);
        if (!lang) {
            s = apply_rot(make_rot(H, dt), s);
        } else {
            const double A = lan.x, B = lan.y, C = lan.z, Dd = lan.w;
            const Rot R = make_rot(H, dth);
            s = apply_rot(R, s);
            L = landau_field(s, A, B, C, Dd, S);
            double3 st = add3(s, mul3(gS * dtq, add3(H, L)));
            L = landau_field(st, A, B, C, Dd, S);
            s = add3(s, mul3(gS * dth, add3(H, L)));
            s = add3(s, mul3(dt, mul3(rs, zq)));
            L = landau_field(s, A, B, C, Dd, S);
            st = add3(s, mul3(gS * dtq, add3(H, L)));
            L = landau_field(st, A, B, C, Dd, S);
            s = add3(s, mul3(gS * dth, add3(H, L)));
            s = apply_rot(R, s);
        }
        prev_s = make_double4(s.x, s.y, s.z, spin_len(s));
        prev_a = a;
        if (j == 0) {
            D.s[a] = prev_s;
            // the fields of the forward sweep are overwritten by the
            // backward sweep before anything reads them
            if (store_fields) {
                if (lang) D.sjs[a] = S;
                D.hh[a] = make_double4(H.x, H.y, H.z, 0.0);
                D.hl[a] = make_double4(L.x, L.y, L.z, 0.0);
            }
        }
    }
}
template __global__ void k_spin_phase<1>(const int*, int, int, int, double, Tables, Phys, Dev);
template __global__ void k_spin_phase<2>(const int*, int, int, int, double, Tables, Phys, Dev);
template __global__ void k_spin_phase<4>(const int*, int, int, int, double, Tables, Phys, Dev);

// ---------------------------------------------------------------------------
// momenta (core_dp_CPU.cpp:51-312). Every part of the Langevin update is local
// to one link cell, so one group of kSeg lanes runs C1, B, A, B, C2 for its
// cell in order. Lane l of a group holds the cell's list entries l, l + kSeg,
// ... (kSegSlots slots; bcc Fe has about six atoms per cell, so most lanes
// hold one): eight cells per wavefront.

constexpr int kSegSlots = kCellCap / kSeg;
static_assert(kCellCap % kSeg == 0, "a cell's list fills whole slots");

// sum over a cell: each lane's slots in order, then a butterfly over the group
__device__ __forceinline__ double seg_sum(const double (&v)[kSegSlots]) {
    double t = v[0];
#pragma unroll
    for (int s = 1; s < kSegSlots; ++s) t += v[s];
#pragma unroll
    for (int m = kSeg / 2; m > 0; m >>= 1) t += __shfl_xor(t, m, kSeg);
    return t;
}

__device__ __forceinline__ double3 seg_sum(const double3 (&v)[kSegSlots]) {
    double x[kSegSlots], y[kSegSlots], z[kSegSlots];
#pragma unroll
    for (int s = 0; s < kSegSlots; ++s) {
        x[s] = v[s].x; y[s] = v[s].y; z[s] = v[s].z;
    }
    return make_double3(seg_sum(x), seg_sum(y), seg_sum(z));
}

// Local collective motion step: walking the list (head to tail, or tail to
// head), each atom gets p += g dt/n (sum_p - p) with sum_p over the cell's
// current momenta (core_dp_CPU.cpp:211-264). The cell sum is carried along:
// after atom k changes by d, the sum changes by d (the reference re-adds the
// list each time; same value up to rounding).
__device__ __forceinline__ void dp_colmot(double3 (&p)[kSegSlots], int l, int n, double factor, bool backward) {
    double3 sp = seg_sum(p);
    const int base = (threadIdx.x & 63) & ~(kSeg - 1);
    for (int q = 0; q < n; ++q) {
        const int k = backward ? n - 1 - q : q;
        const int kl = k % kSeg, ks = k / kSeg;
        double3 d = make_double3(0.0, 0.0, 0.0);
#pragma unroll
        for (int s = 0; s < kSegSlots; ++s) {
            if (l == kl && s == ks) {
                const double3 pn = add3(p[s], mul3(factor, sub3(sp, p[s])));
                d = sub3(pn, p[s]);
                p[s] = pn;
            }
        }
        const int src = base + kl;
        d.x = __shfl(d.x, src, 64);
        d.y = __shfl(d.y, src, 64);
        d.z = __shfl(d.z, src, 64);
        sp = add3(sp, d);
    }
}

// the force of the momentum update: with extforce the reference adds fext to
// the stored force in place first (core_dp_CPU.cpp:69-71,130-132), so a
// second core_dp before the next force evaluation adds it again
__device__ __forceinline__ double3 dp_force(const Phys& P, const Dev& D, int a) {
    double3 f = d3(D.f[a]);
    if (P.extforce) {
        f = add3(f, d3(D.fext[a]));
        D.f[a] = make_double4(f.x, f.y, f.z, 0.0);
    }
    return f;
}

// ---------------------------------------------------------------------------
// quantum thermostat (quantum_noise_CPU.cpp). History element e of an atom's
// three degrees of freedom is normal3 at pass kQnCall0 + e, a range that no
// lattice or spin pass reaches; element e lives in ring slot e % 2Nf.
constexpr uint64_t kQnCall0 = 1ull << 62;

// initial_quantum_noise (:115-119): elements 0 .. 2Nf-1, one thread per atom
__global__ __launch_bounds__(256) void k_qn_init(Phys P, Dev D) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N || D.arep[a] < 0) return;
    const Rep& R = D.rep[D.arep[a]];
    const size_t stride = 3 * (size_t)D.N;
    for (int j = 0; j < P.qnf2; ++j) {
        const double3 z = normal3(R.seed, R.replica, a - R.atom0, kQnCall0 + (uint64_t)j);
        double* m = D.qring + (size_t)j * stride + 3 * (size_t)a;
        m[0] = z.x;
        m[1] = z.y;
        m[2] = z.z;
    }
}

// quantum_noise (:60-81) for the r-th redraw of every degree of freedom at
// once: the noise is sum_i H[i] * element (r + 2Nf - 1 - i), newest first;
// element 2Nf + r then replaces element r, the oldest, in its slot. Each
// lane reads its atom's three doubles of a slot, so a wave reads 1.5 KB
// contiguous per tap.
__global__ __launch_bounds__(256) void k_qn_refresh(Phys P, Dev D, uint64_t r) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N || D.arep[a] < 0) return;
    const int rr = D.arep[a];
    const Rep& R = D.rep[rr];
    const int nf2 = P.qnf2;
    const double* H = D.qH + (size_t)rr * nf2;
    const size_t stride = 3 * (size_t)D.N;
    const double* base = D.qring + 3 * (size_t)a;
    int slot = (int)((r + (uint64_t)nf2 - 1) % (uint64_t)nf2);
    double n0 = 0.0, n1 = 0.0, n2 = 0.0;
    for (int i = 0; i < nf2; ++i) {
        const double* m = base + (size_t)slot * stride;
        const double h = H[i];
        n0 += h * m[0];
        n1 += h * m[1];
        n2 += h * m[2];
        slot = slot == 0 ? nf2 - 1 : slot - 1;
    }
    const double3 z = normal3(R.seed, R.replica, a - R.atom0, kQnCall0 + (uint64_t)nf2 + r);
    double* m = D.qring + (size_t)(r % (uint64_t)nf2) * stride + 3 * (size_t)a;
    m[0] = z.x;
    m[1] = z.y;
    m[2] = z.z;
    double* q = D.qval + 3 * (size_t)a;
    q[0] = n0;
    q[1] = n1;
    q[2] = n2;
}

// One group of kSeg lanes per link cell, eight cells per 64-thread block (a
// replica's cells start on an 8-cell boundary, so each block's partial sums
// belong to one replica).
__global__ __launch_bounds__(64) void k_dp_cells(double dt, Phys P, Dev D, int pass_off, double* partials) {
    const int l = threadIdx.x & (kSeg - 1);
    const int c = (blockIdx.x * blockDim.x + threadIdx.x) / kSeg;
    const int rr = c < D.C ? D.crep[c] : -1;
    const bool valid = rr >= 0;
    const int n = valid ? D.cn[c] : 0;
    bool has[kSegSlots];
    int a[kSegSlots];
    double3 p[kSegSlots];
#pragma unroll
    for (int s = 0; s < kSegSlots; ++s) {
        has[s] = l + kSeg * s < n;
        a[s] = has[s] ? D.cl[c * kCellCap + l + kSeg * s] : 0;
        p[s] = has[s] ? d3(D.p[a[s]]) : make_double3(0.0, 0.0, 0.0);
    }
    if (valid) {
        const double gm = P.gamma_L / P.atmass;
        if (P.lattlang) {
            const double h = dt / 2.0;
            if (P.localcolmot && n > 0) dp_colmot(p, l, n, gm * h / n, false);
            const double eb = P.localcolmot ? exp(-gm * h * (1.0 - 1.0 / n)) : exp(-gm * h);
#pragma unroll
            for (int s = 0; s < kSegSlots; ++s) p[s] = mul3(eb, p[s]);
            const Rep& R = D.rep[rr];
            const double Tk = P.eltemp ? D.Te[c] : R.temperature;
            double3 ff[kSegSlots];
#pragma unroll
            for (int s = 0; s < kSegSlots; ++s) {
                ff[s] = make_double3(0.0, 0.0, 0.0);
                if (has[s]) {
                    if (P.quantumnoise) {
                        // core_dp_CPU.cpp:137-146: coloured noise scaled by sqrt(2 gamma / h), h = M dt
                        const double hq = dt * P.qm;
                        const double len = sqrt(2.0 * P.gamma_L / hq);
                        const double* q = D.qval + 3 * (size_t)a[s];
                        ff[s] = mul3(len, make_double3(q[0], q[1], q[2]));
                    } else {
                        const double len = sqrt(2.0 * Tk * P.gamma_L / dt);
                        const uint64_t pass = D.pass_base[0] + (uint64_t)pass_off;
                        ff[s] = mul3(len, normal3(R.seed, R.replica, a[s] - R.atom0, pass));
                    }
                    p[s] = add3(p[s], mul3(dt, add3(dp_force(P, D, a[s]), ff[s])));
                }
            }
            if (P.localcolmot && n > 0) {
                const double3 fs = seg_sum(ff);
                const double3 ave = make_double3(fs.x / n, fs.y / n, fs.z / n);
#pragma unroll
                for (int s = 0; s < kSegSlots; ++s)
                    if (has[s]) p[s] = sub3(p[s], mul3(dt, ave));
            }
#pragma unroll
            for (int s = 0; s < kSegSlots; ++s) p[s] = mul3(eb, p[s]);
            if (P.localcolmot && n > 0) dp_colmot(p, l, n, gm * h / n, true);
        } else {
#pragma unroll
            for (int s = 0; s < kSegSlots; ++s)
                if (has[s]) p[s] = add3(p[s], mul3(dt, dp_force(P, D, a[s])));
        }
#pragma unroll
        for (int s = 0; s < kSegSlots; ++s)
            if (has[s]) D.p[a[s]] = make_double4(p[s].x, p[s].y, p[s].z, 0.0);
    }
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < kSegSlots; ++s) {
        if (has[s]) {
            acc[0] += p[s].x;
            acc[1] += p[s].y;
            acc[2] += p[s].z;
        }
    }
    block_partials<3>(acc, partials);
}

// rescale_momentum, part 1: remove the mean momentum (core_dp_CPU.cpp:270-302);
// the replica's momentum sum is scal[0..2] of its scalar block
__global__ __launch_bounds__(256) void k_dp_center(Phys P, Dev D, const double* scal, double* partials) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    const int rr = a < D.N ? D.arep[a] : -1;
    if (rr >= 0) {
        const double* psum = scal + (size_t)rr * kScal;
        const double inv_n = 1.0 / D.rep[rr].natom;
        double3 mean = make_double3(psum[0] * inv_n, psum[1] * inv_n, psum[2] * inv_n);
        double3 p = d3(D.p[a]);
        acc[0] = sq3(p) / 2.0 / P.atmass;
        p = sub3(p, mean);
        acc[1] = sq3(p) / 2.0 / P.atmass;
        D.p[a] = make_double4(p.x, p.y, p.z, 0.0);
    }
    block_partials<2>(acc, partials);
}

// rescale_momentum, part 2, and the kinetic energies (core_dp_CPU.cpp:78-84,304-310);
// the replica's kinetic energies before and after centring are scal[3..4]
__global__ __launch_bounds__(256) void k_dp_scale(Phys P, Dev D, const double* scal) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    const int rr = D.arep[a];
    if (rr < 0) return;
    const double* ke2 = scal + (size_t)rr * kScal + 3;
    double factor = sqrt(ke2[0] / ke2[1]);
    if (ke2[1] < 1e-10) factor = 1.0;
    double3 p = mul3(factor, d3(D.p[a]));
    D.p[a] = make_double4(p.x, p.y, p.z, 0.0);
    D.ke[a] = sq3(p) / 2.0 / P.atmass;
}

// core_dr_CPU.cpp:33-44
__global__ __launch_bounds__(256) void k_dr(double dt, Phys P, Dev D) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    const int rr = D.arep[a];
    if (rr < 0) return;
    double k = dt / P.atmass;
    double3 r = add3(d3(D.r[a]), mul3(k, d3(D.p[a])));
    r = wrap(D.rep[rr].box, r);
    D.r[a] = make_double4(r.x, r.y, r.z, 0.0);
}

// ---------------------------------------------------------------------------
// electron subsystem

// calculate_temperature_CPU.cpp:35-148: one group of kSeg lanes per link
// cell, lane l taking the cell's list entries l, l + kSeg, ...; the cell sums
// are each lane's entries in order, then a butterfly over the group.
__device__ __forceinline__ double seg_reduce(double t) {
#pragma unroll
    for (int m = kSeg / 2; m > 0; m >>= 1) t += __shfl_xor(t, m, kSeg);
    return t;
}

__global__ __launch_bounds__(256) void k_temperature(Tables T, Phys P, Dev D) {
    const int l = threadIdx.x & (kSeg - 1);
    const int lb = xcd_block((D.C * kSeg + blockDim.x - 1) / blockDim.x);
    if (lb < 0) return;
    const int c = (lb * blockDim.x + threadIdx.x) / kSeg;
    if (c >= D.C || D.crep[c] < 0) return;
    const int n = D.cn[c];
    double ke = 0.0, rup = 0.0, rdn = 0.0, lup = 0.0, ldn = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for (int k = l; k < n; k += kSeg) {
        const int a = D.cl[c * kCellCap + k];
        ke += D.ke[a];
        const double3 pa = d3(D.p[a]);
        px += pa.x; py += pa.y; pz += pa.z;
        double3 H;
        double S;
        exchange_field(a, D, P, H, S);
        double3 s = d3(D.s[a]);
        rup += sq3(cross3(s, H));
        rdn += dot3(s, H);
        const double4 lan = D.lan[a];
        const double A = lan.x, B = lan.y, C = lan.z, Dd = lan.w;
        double s2 = sq3(s), s4 = s2 * s2;
        double s0 = sqrt(s2);
        double kk = -(2.0 * A + 4.0 * B * s2 + 6.0 * C * s4 + 8.0 * Dd * s4 * s2);
        double hc = -S / s0;
        double3 L = make_double3((kk + hc) * s.x, (kk + hc) * s.y, (kk + hc) * s.z);
        lup += sq3(add3(H, L));
        ldn += 6.0 * A + 20.0 * B * s2 + 42.0 * C * s4 + 72.0 * Dd * s4 * s2 + 2.0 * S / s0;
        // s0 stays local: the reference stores it (calculate_temperature_CPU.cpp:100),
        // but here .w is read by the other cells' lanes of this launch through
        // exchange_field, so writing it would race with those reads. The spin
        // sweep and the uploads already keep .w = spin_len(s).
        D.hh[a] = make_double4(H.x, H.y, H.z, 0.0);
        D.hl[a] = make_double4(L.x, L.y, L.z, 0.0);
        D.sjs[a] = S;
    }
    double sum_ke = seg_reduce(ke);
    double3 ap = make_double3(seg_reduce(px), seg_reduce(py), seg_reduce(pz));
    rup = seg_reduce(rup);
    rdn = seg_reduce(rdn);
    lup = seg_reduce(lup);
    ldn = seg_reduce(ldn);
    if (l != 0) return;
    const Rep& R = D.rep[D.crep[c]];
    const double apc = (double)R.natom / (double)R.ncell;
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
__device__ __forceinline__ double dte_dE(const Mesh& M, int c) {
    double dE = 0.0;
    if (M.lattlang) dE += M.Gel[c] * M.Tl[c];
    if (M.spinlang) dE += M.Ges[c] * M.TsL[c];
    return dE;
}

__device__ __forceinline__ void dte_A(int c, double h, const Mesh& M) {
    double ee = te_to_ee(M.Te[c]);
    ee += h * dte_dE(M, c);
    M.Ee[c] = ee;
    M.Te[c] = ee_to_te(ee);
}

// exp(-G / C(T)) - 1 with C(T) = 3 tanh(bT) (heatcapacity_CPU.cpp:34-57)
__device__ __forceinline__ double relax_factor(double G, double T) {
    return exp(-G / (3.0 * tanh_bt(kCeB * T))) - 1.0;
}

// RK4 of the electron-ion/spin relaxation over h (core_dTe_CPU.cpp:91-129),
// G = (Gel + Ges) h / k_B. The stages are carried as k h (the reference
// divides each k by h and multiplies it back): the same increments up to
// rounding.
__device__ __forceinline__ double dte_rk4(double T0, double G) {
    const double q1 = T0 * relax_factor(G, T0);
    const double T1 = T0 + q1 / 2.0;
    const double q2 = T1 * relax_factor(G, T1);
    const double T2 = T0 + q2 / 2.0;
    const double q3 = T2 * relax_factor(G, T2);
    const double T3 = T0 + q3;
    const double q4 = T3 * relax_factor(G, T3);
    return T0 + (q1 + 2.0 * q2 + 2.0 * q3 + q4) / 6.0;
}

__device__ __forceinline__ double dte_G(int c, double h, const Mesh& M) {
    double G = 0.0;
    if (M.lattlang) G += M.Gel[c];
    if (M.spinlang) G += M.Ges[c];
    return G * (h / kBoltz);
}

__device__ __forceinline__ void dte_B(int c, double h, const Mesh& M) {
    M.Te[c] = dte_rk4(M.Te[c], dte_G(c, h, M));
}

__global__ __launch_bounds__(256) void k_dte_ab(double h, Mesh M, double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (c < M.C) {
        dte_A(c, h, M);
        dte_B(c, h, M);
        double ee = te_to_ee(M.Te[c]);
        M.Ee[c] = ee;
        acc[0] = ee;
    }
    block_partials<1>(acc, partials);
}

// one dependency level of the C diffusion sweep (core_dTe_CPU.cpp:155-197).
// The reference recomputes E(Te) of the cell first; the stored Ee of the
// cell's last update is that value up to one rounding. The six face
// neighbours (map_cells.cpp slots 0/13, 2/14, 12/15) come from the cell index.
// In the backward sweep every cell is updated exactly once, last, so the
// launch also leaves partials of the final energies (partials != null), one
// per wave of 64 level entries: the sums do not depend on the block size.
// the update of one cell of a C-sweep level, given as packed coordinates
// ix | iy << 10 | iz << 20; returns its new Ee
__device__ __forceinline__ double dte_cell_update(int pk, double pref, double idx2, double idy2, double idz2, int nx,
                                                  int ny, int nz, double* Te, double* Ee) {
    const int sl = nx * ny;
    const int ix = pk & 1023, iy = (pk >> 10) & 1023, iz = pk >> 20;
    const int c = ix + iy * nx + iz * sl;
    const int xm = ix == 0 ? nx - 1 : ix - 1, xp = ix + 1 == nx ? 0 : ix + 1;
    const int ym = iy == 0 ? ny - 1 : iy - 1, yp = iy + 1 == ny ? 0 : iy + 1;
    const int zm = iz == 0 ? nz - 1 : iz - 1, zp = iz + 1 == nz ? 0 : iz + 1;
    const int row = iy * nx + iz * sl;
    const double tc = Te[c];
    const double txm = Te[xm + row], txp = Te[xp + row];
    const double tym = Te[ix + ym * nx + iz * sl], typ = Te[ix + yp * nx + iz * sl];
    const double tzm = Te[ix + iy * nx + zm * sl], tzp = Te[ix + iy * nx + zp * sl];
    double ee = Ee[c];
    ee += pref * ((txp + txm - 2.0 * tc) * idx2 + (typ + tym - 2.0 * tc) * idy2 + (tzp + tzm - 2.0 * tc) * idz2);
    Ee[c] = ee;
    Te[c] = ee_to_te(ee);
    return ee;
}

__global__ __launch_bounds__(256) void k_dte_sweep(const int* cells, int ncell, double pref, double idx2,
                                                   double idy2, double idz2, int nx, int ny, int nz, double* Te,
                                                   double* Ee, double* partials) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (t < ncell) acc[0] = dte_cell_update(cells[t], pref, idx2, idy2, idz2, nx, ny, nz, Te, Ee);
    if (partials) {
        const double v = wave_sum(acc[0]);
        if ((threadIdx.x & 63) == 0) partials[t >> 6] = v;
    }
}

// the two totals of a sub-step (see k_reduce2) by one block of 1024 threads
__device__ __forceinline__ void reduce2_block(const double* pa, int na, const double* pb, int nb, double* out) {
    __shared__ double sh[2][16];
    const int t = threadIdx.x;
    double a = 0.0, e = 0.0;
    // both arrays' loads in flight together; each array's terms are added in
    // the order t, t + 1024, ...
    const int nmax = na > nb ? na : nb;
    for (int k0 = t; k0 < nmax; k0 += 8 * 1024) {
        double v[8], w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = k0 + u * 1024 < na ? pa[k0 + u * 1024] : 0.0;
            w[u] = k0 + u * 1024 < nb ? pb[k0 + u * 1024] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a += v[u];
            e += w[u];
        }
    }
    a = wave_sum(a);
    e = wave_sum(e);
    if ((t & 63) == 0) {
        sh[0][t >> 6] = a;
        sh[1][t >> 6] = e;
    }
    __syncthreads();
    if (t == 0) {
        double A = 0.0, E = 0.0;
        for (int w = 0; w < 16; ++w) {
            A += sh[0][w];
            E += sh[1][w];
        }
        out[0] = A;
        out[1] = E;
    }
}

// The trailing small levels of a C sweep in one workgroup of 1024 threads:
// level after level, separated by workgroup barriers (all of the block's
// waves share one compute unit and its write-through L1, so a barrier makes
// one level's stores visible to the next level's loads). Each cell update
// and each backward-sweep partial (one per 64 entries of a level, at
// T.poff[level] + entry / 64) is exactly what a k_dte_sweep launch per level
// would produce. With out != null the block then adds the sub-step's two
// totals from pa[0, na) and pb[0, nb), as k_reduce2 does.
__global__ __launch_bounds__(1024) void k_dte_tail(const int* cells, SweepTail T, double pref, double idx2,
                                                   double idy2, double idz2, int nx, int ny, int nz, double* Te,
                                                   double* Ee, double* partials, const double* pa, int na,
                                                   const double* pb, int nb, double* out) {
    const int t = threadIdx.x;
    for (int l = 0; l < T.nlev; ++l) {
        const int o = T.off[l], n = T.off[l + 1] - o;
        for (int e0 = 0; e0 < n; e0 += 1024) {
            const int e = e0 + t;
            double acc = 0.0;
            if (e < n) acc = dte_cell_update(cells[o + e], pref, idx2, idy2, idz2, nx, ny, nz, Te, Ee);
            if (partials) {
                const double v = wave_sum(acc);
                if ((t & 63) == 0 && e < n) partials[T.poff[l] + (e >> 6)] = v;
            }
        }
        __syncthreads();
    }
    if (out) reduce2_block(pa, na, pb, nb, out);
}

// the two totals of a sub-step: out[0] = sum of the starting-energy partials,
// out[1] = sum of the final-energy partials; one block of 1024 threads in a
// fixed order, independent of how the partials were produced. Each thread
// keeps eight loads of each array in flight, a wave sums by shuffles and
// thread 0 adds the 16 wave sums in order.
__global__ __launch_bounds__(1024) void k_reduce2(const double* pa, int na, const double* pb, int nb,
                                                  double* out) {
    reduce2_block(pa, na, pb, nb, out);
}

// starting and final totals of a sub-step from n partial pairs (xp[0..n):
// start, xp[n..2n): end), summed in a fixed order by every thread
__device__ __forceinline__ double dte_factor(const double* xp, int n) {
    double st = 0.0, en = 0.0;
    for (int k = 0; k < n; ++k) {
        st += xp[k];
        en += xp[n + k];
    }
    return st / en;
}

// end of one sub-step and start of the next: renormalise (core_dTe_CPU.cpp:199-216),
// B(h), A(h), then the next sub-step's A(h), B(h) and E(Te) as its C starts
// In registers: the Te and Ee that the first A(h) leaves are overwritten by
// the second before anything reads them, so only the final values are
// computed and stored (same arithmetic as B, A, A, B applied in turn).
__global__ __launch_bounds__(256) void k_dte_finish_ab(double h, Mesh M, const double* xp, int nxp,
                                                       double* partials) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (c < M.C) {
        const double f = dte_factor(xp, nxp);
        const double G = dte_G(c, h, M);
        const double dE = dte_dE(M, c);
        double T = ee_to_te(M.Ee[c] * f);
        T = dte_rk4(T, G);                  // B(h) ends the sub-step
        double ee = te_to_ee(T) + h * dE;      // A(h) ends it
        ee += h * dE;                          // A(h) starts the next
        T = dte_rk4(ee_to_te(ee), G);       // B(h)
        ee = te_to_ee(T);                      // E(Te) as C starts
        M.Te[c] = T;
        M.Ee[c] = ee;
        acc[0] = ee;
    }
    block_partials<1>(acc, partials);
}

// renormalise to the starting total (core_dTe_CPU.cpp:199-216), then B(h), A(h)
__global__ __launch_bounds__(256) void k_dte_finish(double h, Mesh M, const double* xp, int nxp) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M.C) return;
    double f = dte_factor(xp, nxp);
    double ee = M.Ee[c] * f;
    M.Ee[c] = ee;
    M.Te[c] = ee_to_te(ee);
    dte_B(c, h, M);
    dte_A(c, h, M);
}


// ---------------------------------------------------------------------------
// barostats: positions follow the box (scale_pressure_CPU.cpp:59,
// scale_stress_CPU.cpp:67-80)

// PRESSURE: r = factor r
__global__ __launch_bounds__(256) void k_scale_r(Dev D, double factor) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N || D.arep[a] < 0) return;
    const double4 r = D.r[a];
    D.r[a] = make_double4(factor * r.x, factor * r.y, factor * r.z, r.w);
}

// STRESS: fractional coordinates of the old box, real ones of the new box
__global__ __launch_bounds__(256) void k_rebox_r(Dev D, Box old_box, Box new_box) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N || D.arep[a] < 0) return;
    const double4 r = D.r[a];
    const double* iv = old_box.inv;
    const double* d = new_box.d;
    const double qx = iv[0] * r.x + iv[1] * r.y + iv[3] * r.z;
    const double qy = iv[2] * r.y + iv[4] * r.z;
    const double qz = iv[5] * r.z;
    D.r[a] = make_double4(d[0] * qx + d[1] * qy + d[3] * qz, d[2] * qy + d[4] * qz, d[5] * qz, r.w);
}

// ---------------------------------------------------------------------------
// scale_step (scale_step_CPU.cpp:35-70): the largest |p|^2 and |Heff_H|^2

__global__ __launch_bounds__(256) void k_step_max(Dev D, double* partials) {
    __shared__ double sh[2][256];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = threadIdx.x;
    double pm = 0.0, hm = 0.0;
    if (a < D.N && D.arep[a] >= 0) {
        pm = sq3(d3(D.p[a]));
        hm = sq3(d3(D.hh[a]));
    }
    sh[0][t] = pm;
    sh[1][t] = hm;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) {
            sh[0][t] = fmax(sh[0][t], sh[0][t + s]);
            sh[1][t] = fmax(sh[1][t], sh[1][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        partials[2 * blockIdx.x] = sh[0][0];
        partials[2 * blockIdx.x + 1] = sh[1][0];
    }
}

// out[0..1] = the largest of the pair partials (exact in any order)
__global__ __launch_bounds__(256) void k_reduce_max2(const double* partials, int nblk, double* out) {
    __shared__ double sh[2][256];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int k = t; k < nblk; k += 256) {
        a = fmax(a, partials[2 * k]);
        b = fmax(b, partials[2 * k + 1]);
    }
    sh[0][t] = a;
    sh[1][t] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            sh[0][t] = fmax(sh[0][t], sh[0][t + s]);
            sh[1][t] = fmax(sh[1][t], sh[1][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = sh[0][0];
        out[1] = sh[1][0];
    }
}

// ---------------------------------------------------------------------------
// observables (check_energy_CPU.cpp, check_temperature_CPU.cpp, check_spin_CPU.cpp)

__global__ __launch_bounds__(256) void k_obs_atoms(Tables T, Phys P, Dev D, double* partials,
                                                   int scan_spin) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[kObsAtomK + 1];
#pragma unroll
    for (int k = 0; k <= kObsAtomK; ++k) acc[k] = 0.0;
    if (a < D.N && D.arep[a] >= 0) {
        acc[0] = D.pe[a];
        acc[1] = D.ke[a];
        acc[2] = D.me[a];
        acc[3] = D.me0[a];
        double3 s = d3(D.s[a]);
        acc[4] = s.x; acc[5] = s.y; acc[6] = s.z;
        acc[11] = D.vir[a];
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[12 + k] = D.stress[(size_t)k * D.N + a];
        if (scan_spin) {
            // check_temperature_CPU.cpp:64-77,97-128 (no eltemp)
            double3 H;
            double S;
            exchange_field(a, D, P, H, S);
            acc[7] = sq3(cross3(s, H));
            acc[8] = dot3(s, H);
            const double4 lan = D.lan[a];
            const double A = lan.x, B = lan.y, C = lan.z, Dd = lan.w;
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

// .w = |s| over the atom slots [a0, a0 + n), after the host uploaded spins
__global__ __launch_bounds__(256) void k_spin_len(Dev D, int a0, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    D.s[a0 + k].w = spin_len(d3(D.s[a0 + k]));
}

// magmom: the spins from the moments, s = m / (-g), with |s| in .w
// (core_ds_CPU.cpp:37-45, read_config.cpp:136-139), for atom slots
// [a0, a0 + n)
__global__ __launch_bounds__(256) void k_magmom_s(Dev D, int a0, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double4 m = D.m[a0 + k];
    const double3 s = make_double3(m.x / -kElG, m.y / -kElG, m.z / -kElG);
    D.s[a0 + k] = make_double4(s.x, s.y, s.z, spin_len(s));
}

// magmom: the moments from the spins, m = -g s (core_ds_CPU.cpp:73-81)
__global__ __launch_bounds__(256) void k_magmom_m(Dev D) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    const double4 s = D.s[a];
    D.m[a] = make_double4(-kElG * s.x, -kElG * s.y, -kElG * s.z, 0.0);
}

__global__ __launch_bounds__(256) void k_check_nonfinite(Dev D) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.N) return;
    double4 r = D.r[a], p = D.p[a], s = D.s[a];
    double x = r.x + r.y + r.z + p.x + p.y + p.z + s.x + s.y + s.z;
    if (!isfinite(x)) atomicOr(D.err, kErrNonFinite);
}

}  // namespace sld
