// The differentiable potential model: energies, forces and effective fields
// of frozen configurations as functions of the DD05 and Fe exchange/Landau
// parameters, with their parameter adjoint.
#pragma once

#include <memory>
#include <vector>

namespace sld {

// Parameter vector theta:
//   [0..5]   a_k   density spline coefficients      f(r) = sum a_k (R_k - r)^3, r < R_k   (DD05.cpp:20-43)
//   [6..13]  V_k   pair spline coefficients         V(r) = sum V_k (R_k - r)^3, r > 2      (DD05.cpp:64-164)
//   [14..15] aa bb embedding F(rho)                                                        (DD05.cpp:6-18)
//   [16..17] A rc  exchange J(r) = A (1 - r/rc)^3, r < rc                                 (JijFe.cpp:5-14)
//   [18..21] Landau A B C D (constants for Fe)                                             (JijFe.cpp:18-43)
constexpr int kNParam = 22;
enum ParamIndex { P_A0 = 0, P_V0 = 6, P_AA = 14, P_BB = 15, P_JA = 16, P_JRC = 17, P_LA = 18 };

// Columns of the model tables (10^6 entries each, the reference's sampling
// and 6-point derivative stencil):
//   r value table   [N][8]:  f f' V V' J J' - -
//   r derivative    [N][32]: df/da_k, df'/da_k (k<6, interleaved), dV/dV_k, dV'/dV_k (k<8),
//                            dJ/dA dJ'/dA dJ/drc dJ'/drc
//   rho table       [N][8]:  F F' dF/daa dF'/daa dF/dbb dF'/dbb - -
constexpr int kMRStride = 8;
constexpr int kMRDStride = 32;
constexpr int kMHStride = 8;
enum MRCol { MR_F = 0, MR_DF, MR_V, MR_DV, MR_J, MR_DJ };
enum MHCol { MH_F = 0, MH_DF, MH_AA, MH_DAA, MH_BB, MH_DBB };

struct ModelHostTables {
    double rmax = 0, rhomax = 0, in_over_rmax = 0, in_over_rhomax = 0;
    std::vector<double> r, rd, h;
};

// The reference's parameter values (DD05.cpp, JijFe.cpp).
void model_default_params(double* theta);
// Tables of the model at theta, sampled on [0, rcut_max] and [0, rhomax].
void build_model_tables(const double* theta, double rcut_max, ModelHostTables& t);

}  // namespace sld
