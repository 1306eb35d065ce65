// Tables of the parametric potential model (sld_model.h): the forms of
// DD05.cpp:6-164 and JijFe.cpp:5-43 with their coefficients taken from the
// parameter vector, sampled like potential_CPU.cpp:61-170 (10^6 points, the
// 6-point central-difference derivative). Every parameter except the exchange
// cutoff rc enters a table linearly, so the derivative of a table with respect
// to it is the table of its basis function, built the same way; the rc
// derivative is the table of dJ/drc at the current parameters.
//
// The pair term's quintic bridge (1.8 < r <= 2 A) is evaluated as the bridge
// of the ZBL end alone plus sum_k V_k times the bridge of knot k alone. This
// is the same polynomial as DD05.cpp:88-158's single closed form (the bridge
// is linear in its six end values), but the closed form cancels large terms
// and carries rounding noise of about 1e-9 of V (1e-7 of V' after the
// 6-point derivative) that varies with V; split, the noise sits in the
// V-independent ZBL part and the model is smooth in V_k.
#include <cmath>
#include <thread>
#include <vector>

#include "sld_common.h"
#include "sld_model.h"
#include "sld_tables.h"

namespace sld {
namespace {

// knots of the density and pair splines (DD05.cpp:26-37,70-85)
const double kFR[6] = {3.0, 2.8666666666666666, 2.7333333333333333, 2.6, 2.4, 2.3};
const double kVR[8] = {4.1, 3.8, 3.5, 3.2, 2.9, 2.6, 2.4, 2.3};

template <class F>
double deriv6(F fn, double x) {
    double dx = x / 1000e0;
    double vp = fn(x + dx), vm = fn(x - dx);
    double v2p = fn(x + 2 * dx), v2m = fn(x - 2 * dx);
    double v3p = fn(x + 3 * dx), v3m = fn(x - 3 * dx);
    return (45e0 * (vp - vm) - 9e0 * (v2p - v2m) + (v3p - v3m)) / (60e0 * dx);
}

double cube_knot(double R, double r) { return r < R ? std::pow(R - r, 3) : 0e0; }

double quintic(const Bridge& b, double r) {
    using std::pow;
    const double* a = b.a;
    return a[5] * pow(r, 5) + a[4] * pow(r, 4) + a[3] * pow(r, 3) + a[2] * pow(r, 2) + a[1] * r + a[0];
}

struct Model {
    double a[6], V[8], aa, bb, JA, Jrc;
    Bridge zbridge;           // bridge of the ZBL end alone, zero spline end
    Bridge vbasis[8];         // bridge of each V_k alone, zero ZBL ends
    double rhomax;
};

Model make_model(const double* th) {
    Model m;
    for (int k = 0; k < 6; ++k) m.a[k] = th[P_A0 + k];
    for (int k = 0; k < 8; ++k) m.V[k] = th[P_V0 + k];
    m.aa = th[P_AA];
    m.bb = th[P_BB];
    m.JA = th[P_JA];
    m.Jrc = th[P_JRC];
    const double x2 = 2.0;
    m.zbridge = bridge_from_ends(zbl_ends(), 0e0, 0e0, 0e0);
    const ZblEnds zero = {0e0, 0e0, 0e0};
    for (int k = 0; k < 8; ++k) {
        double g = 0e0, dg = 0e0, d2g = 0e0;
        if (x2 < kVR[k]) {
            g = std::pow(kVR[k] - x2, 3);
            dg = -3e0 * std::pow(kVR[k] - x2, 2);
            d2g = 6e0 * (kVR[k] - x2);
        }
        m.vbasis[k] = bridge_from_ends(zero, g, dg, d2g);
    }
    double f1 = 0e0;
    for (int i = 0; i < 6; ++i) f1 += m.a[i] * cube_knot(kFR[i], 1e0);
    m.rhomax = 2e0 * f1 < 1000e0 ? 1000e0 : 2e0 * f1;
    return m;
}

double f_of(const Model& m, double r) {
    double f = 0e0;
    for (int i = 0; i < 6; ++i)
        if (r < kFR[i]) f += m.a[i] * std::pow(kFR[i] - r, 3);
    return f;
}

double v_of(const Model& m, double r) {
    if (r <= 1.8) return zbl_pair(r);
    if (r <= 2.0) {
        double v = quintic(m.zbridge, r);
        for (int k = 0; k < 8; ++k) v += m.V[k] * quintic(m.vbasis[k], r);
        return v;
    }
    double v = 0e0;
    for (int i = 0; i < 8; ++i)
        if (r < kVR[i]) v += m.V[i] * std::pow((kVR[i] - r), 3);
    return v;
}

// dV/dV_k: the knot term above 2, its share of the bridge between 1.8 and 2
double v_basis(const Model& m, int k, double r) {
    if (r <= 1.8) return 0e0;
    if (r <= 2.0) return quintic(m.vbasis[k], r);
    return cube_knot(kVR[k], r);
}

double j_of(const Model& m, double r) { return r < m.Jrc ? m.JA * std::pow(1.0 - r / m.Jrc, 3) : 0e0; }
double j_basis_a(const Model& m, double r) { return r < m.Jrc ? std::pow(1.0 - r / m.Jrc, 3) : 0e0; }
double j_basis_rc(const Model& m, double r) {
    return r < m.Jrc ? m.JA * 3.0 * std::pow(1.0 - r / m.Jrc, 2) * r / (m.Jrc * m.Jrc) : 0e0;
}

double bigf_of(const Model& m, double rho) {
    if (rho >= 1e0) return -m.aa * std::sqrt(rho);
    return -m.aa * std::sqrt(rho) - m.bb / std::log(2e0) * (1e0 - std::sqrt(rho)) * std::log(2e0 - rho);
}
double bigf_basis_aa(double rho) { return -std::sqrt(rho); }
double bigf_basis_bb(double rho) {
    if (rho >= 1e0) return 0e0;
    return -(1e0 - std::sqrt(rho)) * std::log(2e0 - rho) / std::log(2e0);
}

}  // namespace

void model_default_params(double* th) {
    const double a[6] = {0.9332056681088162, -1.162558782567700, -0.3502026949249225,
                         0.4287820835430028, 4.907925057809273,  -5.307049068415304};
    const double V[8] = {-0.1960674387419232, 0.3687525935422963, -1.505333614924853, 4.948907078156191,
                         -4.894613262753399,  3.468897724782442,  -1.792218099820337, 80.22069592246987};
    for (int k = 0; k < 6; ++k) th[P_A0 + k] = a[k];
    for (int k = 0; k < 8; ++k) th[P_V0 + k] = V[k];
    th[P_AA] = 4.100199340884814;
    th[P_BB] = 1.565647547483517;
    th[P_JA] = 0.870;
    th[P_JRC] = 3.75;
    th[P_LA + 0] = -0.744824;
    th[P_LA + 1] = 0.345295;
    th[P_LA + 2] = -0.00790205;
    th[P_LA + 3] = 0e0;
}

void build_model_tables(const double* theta, double rcut_max, ModelHostTables& t) {
    const Model m = make_model(theta);
    t.rmax = rcut_max;
    t.rhomax = m.rhomax;
    const double finput = (double)kNInput;
    t.in_over_rmax = finput / t.rmax;
    t.in_over_rhomax = finput / t.rhomax;
    t.r.assign((size_t)kNInput * kMRStride, 0.0);
    t.rd.assign((size_t)kNInput * kMRDStride, 0.0);
    t.h.assign((size_t)kNInput * kMHStride, 0.0);
    auto work = [&](int n0, int n1) {
        for (int n = n0; n < n1; ++n) {
            const double r = t.rmax / finput * (double)(n + 1);
            const double rho = t.rhomax / finput * (double)(n + 1);
            double* R = &t.r[(size_t)n * kMRStride];
            R[MR_F] = f_of(m, r);
            R[MR_DF] = deriv6([&](double x) { return f_of(m, x); }, r);
            R[MR_V] = v_of(m, r);
            R[MR_DV] = deriv6([&](double x) { return v_of(m, x); }, r);
            R[MR_J] = j_of(m, r);
            R[MR_DJ] = deriv6([&](double x) { return j_of(m, x); }, r);
            double* D = &t.rd[(size_t)n * kMRDStride];
            for (int k = 0; k < 6; ++k) {
                D[2 * k] = cube_knot(kFR[k], r);
                D[2 * k + 1] = deriv6([&](double x) { return cube_knot(kFR[k], x); }, r);
            }
            for (int k = 0; k < 8; ++k) {
                D[12 + 2 * k] = v_basis(m, k, r);
                D[12 + 2 * k + 1] = deriv6([&](double x) { return v_basis(m, k, x); }, r);
            }
            D[28] = j_basis_a(m, r);
            D[29] = deriv6([&](double x) { return j_basis_a(m, x); }, r);
            D[30] = j_basis_rc(m, r);
            D[31] = deriv6([&](double x) { return j_basis_rc(m, x); }, r);
            double* H = &t.h[(size_t)n * kMHStride];
            H[MH_F] = bigf_of(m, rho);
            H[MH_DF] = deriv6([&](double x) { return bigf_of(m, x); }, rho);
            H[MH_AA] = bigf_basis_aa(rho);
            H[MH_DAA] = deriv6(bigf_basis_aa, rho);
            H[MH_BB] = bigf_basis_bb(rho);
            H[MH_DBB] = deriv6(bigf_basis_bb, rho);
        }
    };
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    std::vector<std::thread> th;
    int chunk = (kNInput + nt - 1) / nt;
    for (unsigned k = 0; k < nt; ++k) {
        int a = k * chunk, b = a + chunk < kNInput ? a + chunk : kNInput;
        if (a < b) th.emplace_back(work, a, b);
    }
    for (auto& x : th) x.join();
}

}  // namespace sld
