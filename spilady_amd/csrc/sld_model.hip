// Kernels of the parametric potential model: forward (energy, forces,
// effective fields of frozen configurations) and the parameter adjoint.
//
// The configurations are the replicas of a context; the neighbour lists are
// the context's (rcut_max, rebuilt whenever positions change). Energies are
// E = sum_i (pe_i + me_i + me0_i), the quantity whose negative gradients are
// the forces and effective fields of calculate_force_energy_CPU.cpp:33-301
// and calculate_spin_CPU.cpp:111-199. Forces and fields are written compact:
// replica r's atoms start at coff[r].
//
// Adjoint: given upstream gradients gE (per replica), gF, gH (per atom),
// grad_theta = sum gE dE/dtheta + gF . dF/dtheta + gH . dH/dtheta. With the
// pair force written as F_i = sum_j (-dudr_ij / r_ij) r_ij, the force term is
// sum over ordered pairs of c_ij d(dudr_ij)/dtheta with
// c_ij = -(gF_i . r_ij) / r_ij; the F'(rho) part of dudr is gathered per atom
// through Q_i = sum_j (c_ij + c_ji) f'(r_ij) (first pass). Pair geometry is
// recomputed in each pass, not stored.
#include <hip/hip_runtime.h>

#include "sld_kernels.h"
#include "sld_model.h"

namespace sld {

template <int NK>
__device__ __forceinline__ void model_partials(const double (&v)[NK], double* partials, int block, int nblocks) {
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
        for (int k = 0; k < NK; ++k) partials[(size_t)k * nblocks + block] = sh[k][0];
    }
}

// density and F'(rho) of every atom (calculate_rho_CPU.cpp:33-92,
// calculate_force_energy_CPU.cpp:98-107; the Fe Landau coefficients do not
// depend on rho)
__global__ __launch_bounds__(256) void k_model_rho(ModelTab M, Phys P, Dev D, double* rho_out, double* fp_out) {
    const int lb = xcd_block((D.N + blockDim.x - 1) / blockDim.x);
    if (lb < 0) return;
    const int i = lb * blockDim.x + threadIdx.x;
    if (i >= D.N) return;
    const int rr = D.arep[i];
    if (rr < 0) return;
    const Box b = D.rep[rr].box;
    const double3 ri = d3(D.r[i]);
    double rho = 0.0;
    const int nn = D.nbn[i];
    const int* nb = D.nb + (size_t)i * kMaxNb;
    for (int k = 0; k < nn; ++k) {
        double rsq;
        (void)min_image_sq(b, sub3(ri, d3(D.r[nb[k]])), rsq);
        if (rsq < P.rcut_pot_sq) {
            int n;
            double ratio;
            interp_index(sqrt(rsq), M.in_over_rmax, n, ratio);
            rho += lerp_tab(M.r, kMRStride, MR_F, n, ratio);
        }
    }
    int n;
    double ratio;
    interp_index(rho, M.in_over_rhomax, n, ratio);
    rho_out[i] = rho;
    fp_out[i] = lerp_tab(M.h, kMHStride, MH_DF, n, ratio);
}

__global__ __launch_bounds__(256) void k_model_forward(ModelTab M, Phys P, Dev D, const double* rho,
                                                       const double* fp, const int* coff, double* F,
                                                       double* H, double* partials) {
    const int lb = xcd_block((D.N + blockDim.x - 1) / blockDim.x);
    if (lb < 0) return;  // whole padding block
    const int i = lb * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    const int rr = i < D.N ? D.arep[i] : -1;
    if (rr >= 0) {
        const Rep& R = D.rep[rr];
        const Box b = R.box;
        const double3 ri = d3(D.r[i]);
        const double3 si = d3(D.s[i]);
        const double si_sq = sq3(si);
        const double fpi = fp[i];
        double3 f = make_double3(0.0, 0.0, 0.0), h = make_double3(0.0, 0.0, 0.0);
        double e = 0.0, S = 0.0;
        const int nn = D.nbn[i];
        const int* nb = D.nb + (size_t)i * kMaxNb;
        for (int k = 0; k < nn; ++k) {
            const int j = nb[k];
            double rsq;
            const double3 rij = min_image_sq(b, sub3(ri, d3(D.r[j])), rsq);
            const double r = sqrt(rsq);
            int n;
            double ratio;
            interp_index(r, M.in_over_rmax, n, ratio);
            const double* t0 = M.r + (size_t)(n - 1) * kMRStride;
            const double* t1 = t0 + kMRStride;
            const double wr = 1.0 - ratio;
            double dudr = 0.0;
            if (r < P.rcut_pot) {
                dudr = (fpi + fp[j]) * (t0[MR_DF] * ratio + t1[MR_DF] * wr) + (t0[MR_DV] * ratio + t1[MR_DV] * wr);
                e += 0.5 * (t0[MR_V] * ratio + t1[MR_V] * wr);
            }
            if (r < P.rcut_mag) {
                const double4 sj4 = D.s[j];
                const double3 sj = d3(sj4);
                const double X = dot3(si, sj) - sqrt(si_sq * sq3(sj));
                const double J = t0[MR_J] * ratio + t1[MR_J] * wr;
                dudr += -(t0[MR_DJ] * ratio + t1[MR_DJ] * wr) * X;
                e += -0.5 * J * X;
                h = add3(h, mul3(J, sj));
                S += J * sj4.w;
            }
            f = add3(f, mul3(-dudr / r, rij));
        }
        int n;
        double ratio;
        interp_index(rho[i], M.in_over_rhomax, n, ratio);
        const double s2 = si_sq, s4 = s2 * s2;
        e += lerp_tab(M.h, kMHStride, MH_F, n, ratio);
        e += M.L[0] * s2 + M.L[1] * s4 + M.L[2] * s4 * s2 + M.L[3] * s4 * s4;
        const double kl = -(2.0 * M.L[0] + 4.0 * M.L[1] * s2 + 6.0 * M.L[2] * s4 + 8.0 * M.L[3] * s4 * s2);
        const double s0 = sqrt(s2);
        const double hc = s0 > 0.0 ? -S / s0 : 0.0;
        h = add3(h, mul3(kl + hc, si));
        const size_t c = (size_t)coff[rr] + (i - R.atom0);
        F[3 * c] = f.x; F[3 * c + 1] = f.y; F[3 * c + 2] = f.z;
        H[3 * c] = h.x; H[3 * c + 1] = h.y; H[3 * c + 2] = h.z;
        acc[0] = e;
    }
    model_partials<1>(acc, partials, lb, (D.N + blockDim.x - 1) / blockDim.x);
}

// adjoint, first pass: Q_i = sum_j (c_ij + c_ji) f'(r_ij)
__global__ __launch_bounds__(256) void k_model_q(ModelTab M, Phys P, Dev D, const int* coff, const double* gF,
                                                 double* Q) {
    const int lb = xcd_block((D.N + blockDim.x - 1) / blockDim.x);
    if (lb < 0) return;
    const int i = lb * blockDim.x + threadIdx.x;
    if (i >= D.N) return;
    const int rr = D.arep[i];
    if (rr < 0) return;
    const Rep& R = D.rep[rr];
    const Box b = R.box;
    const double3 ri = d3(D.r[i]);
    const size_t ci = (size_t)coff[rr] + (i - R.atom0);
    const double3 gi = make_double3(gF[3 * ci], gF[3 * ci + 1], gF[3 * ci + 2]);
    double q = 0.0;
    const int nn = D.nbn[i];
    const int* nb = D.nb + (size_t)i * kMaxNb;
    for (int k = 0; k < nn; ++k) {
        const int j = nb[k];
        double rsq;
        const double3 rij = min_image_sq(b, sub3(ri, d3(D.r[j])), rsq);
        if (rsq >= P.rcut_pot_sq) continue;
        const double r = sqrt(rsq);
        const size_t cj = (size_t)coff[rr] + (j - R.atom0);
        const double3 gj = make_double3(gF[3 * cj], gF[3 * cj + 1], gF[3 * cj + 2]);
        int n;
        double ratio;
        interp_index(r, M.in_over_rmax, n, ratio);
        q += -dot3(sub3(gi, gj), rij) / r * lerp_tab(M.r, kMRStride, MR_DF, n, ratio);
    }
    Q[i] = q;
}

// adjoint, second pass: the kNParam components of grad_theta per atom
__global__ __launch_bounds__(256) void k_model_adjoint(ModelTab M, Phys P, Dev D, const double* rho,
                                                       const double* fp, const double* Q, const int* coff,
                                                       const double* gE, const double* gF, const double* gH,
                                                       double* partials) {
    const int lb = xcd_block((D.N + blockDim.x - 1) / blockDim.x);
    if (lb < 0) return;  // whole padding block
    const int i = lb * blockDim.x + threadIdx.x;
    double acc[kNParam];
#pragma unroll
    for (int k = 0; k < kNParam; ++k) acc[k] = 0.0;
    const int rr = i < D.N ? D.arep[i] : -1;
    if (rr >= 0) {
        const Rep& R = D.rep[rr];
        const Box b = R.box;
        const double3 ri = d3(D.r[i]);
        const double3 si = d3(D.s[i]);
        const double si_sq = sq3(si), si0 = sqrt(si_sq);
        const double fpi = fp[i];
        const size_t ci = (size_t)coff[rr] + (i - R.atom0);
        const double ge = gE[rr];
        const double3 gf = make_double3(gF[3 * ci], gF[3 * ci + 1], gF[3 * ci + 2]);
        const double3 gh = make_double3(gH[3 * ci], gH[3 * ci + 1], gH[3 * ci + 2]);
        const double gh_si = dot3(gh, si);
        double drho[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const int nn = D.nbn[i];
        const int* nb = D.nb + (size_t)i * kMaxNb;
        for (int k = 0; k < nn; ++k) {
            const int j = nb[k];
            double rsq;
            const double3 rij = min_image_sq(b, sub3(ri, d3(D.r[j])), rsq);
            const double r = sqrt(rsq);
            const double cij = -dot3(gf, rij) / r;
            int n;
            double ratio;
            interp_index(r, M.in_over_rmax, n, ratio);
            const double* d0 = M.rd + (size_t)(n - 1) * kMRDStride;
            const double* d1 = d0 + kMRDStride;
            const double wr = 1.0 - ratio;
            if (r < P.rcut_pot) {
                const double fsum = fpi + fp[j];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    drho[q] += d0[2 * q] * ratio + d1[2 * q] * wr;
                    acc[P_A0 + q] += cij * fsum * (d0[2 * q + 1] * ratio + d1[2 * q + 1] * wr);
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const double bv = d0[12 + 2 * q] * ratio + d1[12 + 2 * q] * wr;
                    const double bdv = d0[13 + 2 * q] * ratio + d1[13 + 2 * q] * wr;
                    acc[P_V0 + q] += ge * 0.5 * bv + cij * bdv;
                }
            }
            if (r < P.rcut_mag) {
                const double4 sj4 = D.s[j];
                const double3 sj = d3(sj4);
                const double X = dot3(si, sj) - sqrt(si_sq * sq3(sj));
                const double hv = dot3(gh, sj) - (si0 > 0.0 ? sj4.w / si0 * gh_si : 0.0);
                const double ba = d0[28] * ratio + d1[28] * wr, bda = d0[29] * ratio + d1[29] * wr;
                const double bc = d0[30] * ratio + d1[30] * wr, bdc = d0[31] * ratio + d1[31] * wr;
                acc[P_JA] += ge * (-0.5 * ba * X) + cij * (-bda * X) + ba * hv;
                acc[P_JRC] += ge * (-0.5 * bc * X) + cij * (-bdc * X) + bc * hv;
            }
        }
        // one-body terms: F(rho_i) (energy), F'(rho_i) (through Q_i), Landau
        int n;
        double ratio;
        interp_index(rho[i], M.in_over_rhomax, n, ratio);
        const double* h0 = M.h + (size_t)(n - 1) * kMHStride;
        const double* h1 = h0 + kMHStride;
        const double slope_f = (h1[MH_F] - h0[MH_F]) * M.in_over_rhomax;
        const double slope_fp = (h1[MH_DF] - h0[MH_DF]) * M.in_over_rhomax;
        const double qi = Q[i];
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[P_A0 + q] += (ge * slope_f + qi * slope_fp) * drho[q];
        acc[P_AA] += ge * lerp_tab(M.h, kMHStride, MH_AA, n, ratio) + qi * lerp_tab(M.h, kMHStride, MH_DAA, n, ratio);
        acc[P_BB] += ge * lerp_tab(M.h, kMHStride, MH_BB, n, ratio) + qi * lerp_tab(M.h, kMHStride, MH_DBB, n, ratio);
        const double s2 = si_sq, s4 = s2 * s2, s6 = s4 * s2;
        acc[P_LA + 0] += ge * s2 - 2.0 * gh_si;
        acc[P_LA + 1] += ge * s4 - 4.0 * s2 * gh_si;
        acc[P_LA + 2] += ge * s6 - 6.0 * s4 * gh_si;
        acc[P_LA + 3] += ge * s4 * s4 - 8.0 * s6 * gh_si;
    }
    model_partials<kNParam>(acc, partials, lb, (D.N + blockDim.x - 1) / blockDim.x);
}

}  // namespace sld
