// Host-side construction of the potential tables: the DD05 magnetic EAM of
// Dudarev and Derlet (DD05.cpp:6-164), the Fe exchange J(r) and Landau
// coefficients (JijFe.cpp:5-43), sampled at 10^6 points with a 6-point
// central-difference derivative (potential_CPU.cpp:61-170).
//
// Built with -ffp-contract=off and the same libm calls as the reference so the
// tables agree with the reference's entry for entry.
#include <cmath>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "sld_common.h"
#include "sld_tables.h"

namespace sld {
namespace {

double bigf_gen(double rho) {
    const double aa = 4.100199340884814;
    const double bb = 1.565647547483517;
    if (rho >= 1e0) return -aa * std::sqrt(rho);
    return -aa * std::sqrt(rho) - bb / std::log(2e0) * (1e0 - std::sqrt(rho)) * std::log(2e0 - rho);
}

const double kSfA[6] = {0.9332056681088162, -1.162558782567700, -0.3502026949249225,
                        0.4287820835430028, 4.907925057809273,  -5.307049068415304};
const double kSfR[6] = {3.0, 2.8666666666666666, 2.7333333333333333, 2.6, 2.4, 2.3};

double smallf_gen(double r) {
    double f = 0e0;
    for (int i = 0; i < 6; ++i)
        if (r < kSfR[i]) f += kSfA[i] * std::pow(kSfR[i] - r, 3);
    return f;
}

}  // namespace

double zbl_pair(double r) {
    const double b0 = 0.1818e0, b1 = 0.5099e0, b2 = 0.2802e0, b3 = 0.02817e0;
    const double z = 26e0;
    const double abohr = 0.52917720859e0;
    const double as = 0.88535e0 * abohr / (2e0 * std::pow(z, 0.23));
    const double x = r / as;
    double v = b0 * std::exp(-3.2 * x) + b1 * std::exp(-0.9423 * x) + b2 * std::exp(-0.4029 * x) +
               b3 * std::exp(-0.2016 * x);
    return z * z * 14.3992 * v / r;
}

namespace {

double zbl(double r) { return zbl_pair(r); }

const double kPV[8] = {-0.1960674387419232, 0.3687525935422963, -1.505333614924853,
                       4.948907078156191,   -4.894613262753399, 3.468897724782442,
                       -1.792218099820337,  80.22069592246987};
const double kPR[8] = {4.1, 3.8, 3.5, 3.2, 2.9, 2.6, 2.4, 2.3};

}  // namespace

// Quintic bridge between ZBL at x1 = 1.8 and the knot spline at x2 = 2.0,
// matching value, slope and curvature at both ends (DD05.cpp:93-150): the
// closed-form solution of that 6x6 Hermite system. It is linear in the six
// end values.
ZblEnds zbl_ends() {
    using std::pow;
    const double x1 = 1.8;
    ZblEnds e;
    e.y1 = zbl_pair(x1);
    e.Dy1 = (zbl_pair(x1 + 1e-7) - zbl_pair(x1 - 1e-7)) / 2e-7;
    e.D2y1 = (zbl_pair(x1 + 1e-7) + zbl_pair(x1 - 1e-7) - 2e0 * zbl_pair(x1)) / pow(1e-7, 2);
    return e;
}

Bridge bridge_from_ends(const ZblEnds& e, double y2, double Dy2, double D2y2) {
    using std::pow;
    const double x1 = 1.8, x2 = 2.0;
    const double y1 = e.y1, Dy1 = e.Dy1, D2y1 = e.D2y1;
    const double den = 2e0 * pow((x1 - x2), 5);
    Bridge b;
    b.a[5] = (-6e0*Dy1*x1-6e0*Dy2*x1+D2y1*pow(x1,2)-D2y2*pow(x1,2)
             +6e0*Dy1*x2+6e0*Dy2*x2-2e0*D2y1*x1*x2+2e0*D2y2*x1*x2
             +D2y1*pow(x2,2)-D2y2*pow(x2,2)+12*y1-12*y2) / den;
    b.a[4] = (-2e0*D2y1*pow(x1,3)+3e0*D2y2*pow(x1,3)+D2y1*pow(x1,2)*x2
             -4e0*D2y2*pow(x1,2)*x2+4*D2y1*x1*pow(x2,2)-D2y2*x1*pow(x2,2)
             -3e0*D2y1*pow(x2,3)+2e0*D2y2*pow(x2,3)
             +2e0*Dy1*(7e0*pow(x1,2)+x1*x2-8e0*pow(x2,2))
             +2e0*Dy2*(8e0*pow(x1,2)-x1*x2-7e0*pow(x2,2))
             -30e0*x1*y1-30e0*x2*y1+30e0*x1*y2+30e0*x2*y2) / den;
    b.a[3] = -(-D2y1*pow(x1,4)+3e0*D2y2*pow(x1,4)-4e0*D2y1*pow(x1,3)*x2
             +8e0*D2y1*pow(x1,2)*pow(x2,2)-8e0*D2y2*pow(x1,2)*pow(x2,2)
             +4e0*D2y2*x1*pow(x2,3)-3e0*D2y1*pow(x2,4)+D2y2*pow(x2,4)
             +4e0*Dy1*(2e0*pow(x1,3)+8e0*pow(x1,2)*x2-7e0*x1*pow(x2,2)-3e0*pow(x2,3))
             +4e0*Dy2*(3e0*pow(x1,3)+7e0*pow(x1,2)*x2-8e0*x1*pow(x2,2)-2e0*pow(x2,3))
             -20e0*pow(x1,2)*y1-80e0*x1*x2*y1-20e0*pow(x2,2)*y1
             +20e0*pow(x1,2)*y2+80e0*x1*x2*y2+20e0*pow(x2,2)*y2) / den;
    b.a[2] = (D2y2*x1*pow((x1-x2),2)*(pow(x1,2)+6e0*x1*x2+3e0*pow(x2,2))
             -x2*(3e0*D2y1*pow(x1,4)-8e0*D2y1*pow(x1,2)*pow(x2,2)
             +4e0*D2y1*x1*pow(x2,3)+D2y1*pow(x2,4)
             -12e0*Dy1*x1*(2e0*pow(x1,2)+x1*x2-3e0*pow(x2,2))
             +12e0*Dy2*x1*(-3e0*pow(x1,2)+x1*x2+2e0*pow(x2,2))
             +60e0*pow(x1,2)*y1+60e0*x1*x2*y1
             -60e0*pow(x1,2)*y2-60e0*x1*x2*y2)) / den;
    b.a[1] = (2e0*Dy2*pow(x1,2)*(pow(x1,3)-5e0*pow(x1,2)*x2-8e0*x1*pow(x2,2)+12e0*pow(x2,3))
             +x2*(-D2y2*pow(x1,2)*pow((x1-x2),2)*(2e0*x1+3e0*x2)
             +x2*(-2e0*Dy1*(12e0*pow(x1,3)-8e0*pow(x1,2)*x2-5e0*x1*pow(x2,2)+pow(x2,3))
             +x1*(D2y1*pow((x1-x2),2)*(3e0*x1+2e0*x2)+60e0*x1*(y1-y2))))) / den;
    b.a[0] = (D2y2*pow(x1,3)*pow((x1-x2),2)*pow(x2,2)+8e0*Dy1*pow(x1,3)*pow(x2,3)
             -D2y1*pow(x1,4)*pow(x2,3)-10e0*Dy1*pow(x1,2)*pow(x2,4)
             +2e0*D2y1*pow(x1,3)*pow(x2,4)+2e0*Dy1*x1*pow(x2,5)-D2y1*pow(x1,2)*pow(x2,5)
             -2e0*Dy2*pow(x1,3)*x2*(pow(x1,2)-5e0*x1*x2+4e0*pow(x2,2))
             -20e0*pow(x1,2)*pow(x2,3)*y1+10e0*x1*pow(x2,4)*y1-2e0*pow(x2,5)*y1
             +2e0*pow(x1,5)*y2-10e0*pow(x1,4)*x2*y2+20e0*pow(x1,3)*pow(x2,2)*y2) / den;
    return b;
}


namespace {

Bridge make_bridge() {
    using std::pow;
    const double x2 = 2.0;
    double y2 = 0e0, Dy2 = 0e0, D2y2 = 0e0;
    for (int i = 0; i < 8; ++i) if (x2 < kPR[i]) y2 += kPV[i] * pow((kPR[i] - x2), 3);
    for (int i = 0; i < 8; ++i) if (x2 < kPR[i]) Dy2 -= 3e0 * kPV[i] * pow((kPR[i] - x2), 2);
    for (int i = 0; i < 8; ++i) if (x2 < kPR[i]) D2y2 += 6e0 * kPV[i] * (kPR[i] - x2);
    return bridge_from_ends(zbl_ends(), y2, Dy2, D2y2);
}

const Bridge& bridge() {
    static const Bridge b = make_bridge();
    return b;
}


double pair_gen(double r) {
    using std::pow;
    if (r <= 1.8) return zbl(r) + 0e0;
    if (r <= 2.0) {
        const double* a = bridge().a;
        return a[5]*pow(r,5)+a[4]*pow(r,4)+a[3]*pow(r,3)+a[2]*pow(r,2)+a[1]*r+a[0];
    }
    double v = 0e0;
    for (int i = 0; i < 8; ++i)
        if (r < kPR[i]) v += kPV[i] * pow((kPR[i] - r), 3);
    return v;
}

double jij_gen(double r) {
    const double A = 0.870, rc = 3.75;
    return r < rc ? A * std::pow(1.0 - r / rc, 3) : 0e0;
}
double landauA(double) { return -0.744824; }
double landauB(double) { return 0.345295; }
double landauC(double) { return -0.00790205; }
double landauD(double) { return 0e0; }

template <class F>
double deriv6(F fn, double x) {
    double dx = x / 1000e0;
    double vp = fn(x + dx), vm = fn(x - dx);
    double v2p = fn(x + 2 * dx), v2m = fn(x - 2 * dx);
    double v3p = fn(x + 3 * dx), v3m = fn(x - 3 * dx);
    return (45e0 * (vp - vm) - 9e0 * (v2p - v2m) + (v3p - v3m)) / (60e0 * dx);
}

}  // namespace

void build_tables(double rcut_max, bool magmom, HostTables& t) {
    t.rmax = rcut_max;
    t.rhomax = 2e0 * smallf_gen(1e0);
    if (t.rhomax < 1000e0) t.rhomax = 1000e0;
    const double finput = (double)kNInput;
    t.in_over_rmax = finput / t.rmax;
    t.in_over_rhomax = finput / t.rhomax;
    t.rtab.assign((size_t)kNInput * kRStride, 0.0);
    t.rhotab.assign((size_t)kNInput * kRhoStride, 0.0);
    (void)bridge();
    auto work = [&](int n0, int n1) {
        for (int n = n0; n < n1; ++n) {
            double r = t.rmax / (double)kNInput * (double)(n + 1);
            double rho = t.rhomax / (double)kNInput * (double)(n + 1);
            double* R = &t.rtab[(size_t)n * kRStride];
            R[R_SF] = smallf_gen(r);
            R[R_DSF] = deriv6(smallf_gen, r);
            R[R_PR] = pair_gen(r);
            R[R_DPR] = deriv6(pair_gen, r);
            R[R_J] = jij_gen(r);
            R[R_DJ] = deriv6(jij_gen, r);
            double* H = &t.rhotab[(size_t)n * kRhoStride];
            H[H_F] = bigf_gen(rho);
            H[H_DF] = deriv6(bigf_gen, rho);
            H[H_A] = landauA(rho);
            H[H_B] = landauB(rho);
            H[H_C] = landauC(rho);
            H[H_D] = landauD(rho);
            H[H_DA] = deriv6(landauA, rho);
            H[H_DB] = deriv6(landauB, rho);
            H[H_DC] = deriv6(landauC, rho);
            H[H_DD] = deriv6(landauD, rho);
            if (magmom) {
                R[R_J] *= std::pow(kElG, 2);
                R[R_DJ] *= std::pow(kElG, 2);
                H[H_A] *= std::pow(kElG, 2);
                H[H_B] *= std::pow(kElG, 4);
                H[H_C] *= std::pow(kElG, 6);
                H[H_D] *= std::pow(kElG, 8);
                H[H_DA] *= std::pow(kElG, 2);
                H[H_DB] *= std::pow(kElG, 4);
                H[H_DC] *= std::pow(kElG, 6);
                H[H_DD] *= std::pow(kElG, 8);
            }
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
    // interpolation rows of the device r table
    t.prow.assign((size_t)kNInput * kPRowStride, 0.0);
    const int cols[6][2] = {{P_DSF, R_DSF}, {P_DPR, R_DPR}, {P_PR, R_PR}, {P_J, R_J}, {P_DJ, R_DJ}, {P_SF, R_SF}};
    for (int n = 1; n < kNInput; ++n) {
        double* q = &t.prow[(size_t)n * kPRowStride];
        for (const auto& cc : cols) {
            q[cc[0]] = t.rtab[(size_t)(n - 1) * kRStride + cc[1]];
            q[cc[0] + 1] = t.rtab[(size_t)n * kRStride + cc[1]];
        }
    }
}

std::shared_ptr<const HostTables> shared_tables(double rcut_max, bool magmom) {
    static std::mutex mu;
    struct Entry {
        double rcut_max;
        bool magmom;
        std::shared_ptr<const HostTables> t;
    };
    static std::vector<Entry> cache;
    std::lock_guard<std::mutex> lock(mu);
    for (auto& e : cache)
        if (e.rcut_max == rcut_max && e.magmom == magmom) return e.t;
    auto t = std::make_shared<HostTables>();
    build_tables(rcut_max, magmom, *t);
    cache.push_back({rcut_max, magmom, t});
    return t;
}

}  // namespace sld
