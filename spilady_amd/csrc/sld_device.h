// Device-side helpers of the spin-lattice engine (gfx950, FP64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "heatcap_poly.h"
#include "sld_common.h"

// Floating-point contraction is decided by the source alone: a multiply and
// an add fuse only inside one expression, never across statements by the
// backend. One device function inlined into two kernels (or two template
// instances of one kernel) then rounds identically in both, which the
// bitwise equalities between launch shapes, batched and single contexts, and
// graph and eager launches rest on.
#pragma clang fp contract(on)

namespace sld {

// ---------------------------------------------------------------------------
// small vector helpers on double4 (w is a payload, never part of the algebra)
__device__ __forceinline__ double3 d3(const double4& a) { return make_double3(a.x, a.y, a.z); }
__device__ __forceinline__ double3 add3(double3 a, double3 b) { return make_double3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ double3 sub3(double3 a, double3 b) { return make_double3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ double3 mul3(double k, double3 a) { return make_double3(k * a.x, k * a.y, k * a.z); }
__device__ __forceinline__ double dot3(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double sq3(double3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// |s|, the value every spin's .w payload holds. It is the only definition of
// that value: the field sums read a neighbour's .w where the reference takes
// vec_length(s_j) (calculate_spin_CPU.cpp:265), so every writer of .w must
// round it the same way. The explicit fma order leaves the compiler no
// contraction choice that could differ between kernels.
__device__ __forceinline__ double spin_len(double3 s) { return sqrt(fma(s.z, s.z, fma(s.x, s.x, s.y * s.y))); }
__device__ __forceinline__ double3 cross3(double3 a, double3 b) {
    return make_double3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}

// ---------------------------------------------------------------------------
// Box geometry (upper-triangular box matrix, find_image_CPU.cpp:31-53 and
// periodic_CPU.cpp:31-50).
struct Box {
    double d[6];    // xx yx yy zx zy zz
    double inv[6];  // inverse, same layout
    double safe_sq; // below this squared length a vector's fractional coordinates all lie in (-1/2, 1/2)
    int nx, ny, nz;
};

__device__ __forceinline__ double3 to_frac(const Box& b, double3 r) {
    return make_double3(b.inv[0] * r.x + b.inv[1] * r.y + b.inv[3] * r.z,
                        b.inv[2] * r.y + b.inv[4] * r.z, b.inv[5] * r.z);
}

// minimum image: one +-1 shift per axis; the vector is rebuilt from q only
// when a shift happened.
__device__ __forceinline__ double3 min_image(const Box& b, double3 rij) {
    double3 q = to_frac(b, rij);
    int idx = 0;
    if (q.x < -0.5) { q.x += 1.0; ++idx; }
    if (q.x >= 0.5) { q.x -= 1.0; ++idx; }
    if (q.y < -0.5) { q.y += 1.0; ++idx; }
    if (q.y >= 0.5) { q.y -= 1.0; ++idx; }
    if (q.z < -0.5) { q.z += 1.0; ++idx; }
    if (q.z >= 0.5) { q.z -= 1.0; ++idx; }
    if (idx > 0) {
        rij.x = b.d[0] * q.x + b.d[1] * q.y + b.d[3] * q.z;
        rij.y = b.d[2] * q.y + b.d[4] * q.z;
        rij.z = b.d[5] * q.z;
    }
    return rij;
}

// the minimum image of rij and its squared length: a vector shorter than
// sqrt(b.safe_sq) needs no shift, so min_image (and its arithmetic) is
// evaluated only past that length, with the same result
__device__ __forceinline__ double3 min_image_sq(const Box& b, double3 rij, double& rsq) {
    rsq = sq3(rij);
    if (rsq >= b.safe_sq) {
        rij = min_image(b, rij);
        rsq = sq3(rij);
    }
    return rij;
}

// wrap into the box; q in (-1, 0) is left as is (int() truncates toward 0),
// exactly as the reference does.
__device__ __forceinline__ double3 wrap(const Box& b, double3 r) {
    double3 q = to_frac(b, r);
    int idx = 0;
    if (q.x < 0.0 || q.x >= 1.0) { q.x -= (double)(int)q.x; ++idx; }
    if (q.y < 0.0 || q.y >= 1.0) { q.y -= (double)(int)q.y; ++idx; }
    if (q.z < 0.0 || q.z >= 1.0) { q.z -= (double)(int)q.z; ++idx; }
    if (idx > 0) {
        r.x = b.d[0] * q.x + b.d[1] * q.y + b.d[3] * q.z;
        r.y = b.d[2] * q.y + b.d[4] * q.z;
        r.z = b.d[5] * q.z;
    }
    return r;
}

__device__ __forceinline__ int bin_cell(const Box& b, double3 r) {
    double3 q = to_frac(b, r);
    int lx = (int)(q.x * b.nx), ly = (int)(q.y * b.ny), lz = (int)(q.z * b.nz);
    lx = lx < 0 ? 0 : (lx >= b.nx ? b.nx - 1 : lx);
    ly = ly < 0 ? 0 : (ly >= b.ny ? b.ny - 1 : ly);
    lz = lz < 0 ? 0 : (lz >= b.nz ? b.nz - 1 : lz);
    return lx + ly * b.nx + lz * b.nx * b.ny;
}

// The 26 neighbours of a link cell in the reference's slot order
// (map_cells.cpp:96-121): slot -> (dx, dy, dz).
__constant__ static const signed char kNeighOff[26][3] = {
    {1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0},  {1, 0, -1}, {1, 1, -1}, {0, 1, -1},
    {-1, 1, -1}, {1, 0, 1}, {1, 1, 1},  {0, 1, 1},   {-1, 1, 1}, {0, 0, 1},  {-1, 0, 0},
    {0, -1, 0}, {0, 0, -1}, {1, -1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1}, {0, -1, 1},
    {0, -1, -1}, {1, -1, 1}, {1, -1, -1}, {-1, -1, 1}, {-1, -1, -1}};

__device__ __forceinline__ int neigh_cell(const Box& b, int c, int slot) {
    int ix = c % b.nx, iy = (c / b.nx) % b.ny, iz = c / (b.nx * b.ny);
    int x = ix + kNeighOff[slot][0], y = iy + kNeighOff[slot][1], z = iz + kNeighOff[slot][2];
    x = x < 0 ? x + b.nx : (x >= b.nx ? x - b.nx : x);
    y = y < 0 ? y + b.ny : (y >= b.ny ? y - b.ny : y);
    z = z < 0 ? z + b.nz : (z >= b.nz ? z - b.nz : z);
    return x + y * b.nx + z * b.nx * b.ny;
}

// ---------------------------------------------------------------------------
// Tables: linear interpolation between entries n-1 and n with
// n = int(x * N / max) (potential_CPU.cpp:172-179). Entry k holds the value
// at x = max * (k+1) / N.
//   r tables interleaved per entry: [sf dsf pr dpr J dJ - -]
//   rho tables interleaved per entry: [F dF A B C D dA dB dC dD - ...]
struct Tables {
    const double* prow;    // [NINPUT][16]: interpolation pairs of the r columns
    const double* rhotab;  // [NINPUT][16]
    double in_over_rmax, in_over_rhomax;
};

// Workgroups are dealt to the eight XCDs round-robin: block b runs on XCD
// b % 8, and each XCD has its own L2. xcd_block maps a launched block to a
// logical block so that every XCD works through one contiguous range of
// logical blocks; spatially neighbouring atoms (whose positions, spins and
// table rows a pair loop reads) then meet in the same L2. The grid is
// launched rounded up to a multiple of 8; padding blocks get -1.
__device__ __forceinline__ int xcd_block(int nlogical) {
    const int b = blockIdx.x;
    const int per = (nlogical + 7) >> 3;
    const int L = (b & 7) * per + (b >> 3);
    return L < nlogical ? L : -1;
}

__device__ __forceinline__ void interp_index(double x, double in_over_max, int& n, double& ratio) {
    double fn = x * in_over_max;
    n = (int)fn;
    // the reference reads entry n-1 with n = int(x N / max); x below one
    // table step or at/above max (or not finite) would leave the table, so n
    // is kept in [1, N-1]
    if (!(n >= 1)) n = 1;
    if (n > kNInput - 1) n = kNInput - 1;
    ratio = (double)(n + 1) - fn;
}

__device__ __forceinline__ double lerp_tab(const double* tab, int stride, int k, int n, double ratio) {
    return tab[(size_t)(n - 1) * stride + k] * ratio + tab[(size_t)n * stride + k] * (1.0 - ratio);
}

// ---------------------------------------------------------------------------
// Counter-based noise (Philox4x32-10, Salmon et al. SC'11): key = (seed low,
// seed high ^ replica), counter = (atom, pass low, pass high, 'SPLD'); four
// words -> uniforms in (0,1) -> two Box-Muller pairs, first three normals used.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double3 normal3(uint64_t seed, int replica, int atom, uint64_t pass) {
    uint32_t c[4] = {(uint32_t)atom, (uint32_t)pass, (uint32_t)(pass >> 32), 0x53504C44u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32) ^ (uint32_t)replica);
    const double inv32 = 2.3283064365386963e-10;
    double u0 = ((double)c[0] + 0.5) * inv32;
    double u1 = ((double)c[1] + 0.5) * inv32;
    double u2 = ((double)c[2] + 0.5) * inv32;
    double u3 = ((double)c[3] + 0.5) * inv32;
    const double two_pi = 6.283185307179586;
    double a = sqrt(-2.0 * log(u0));
    double b = sqrt(-2.0 * log(u2));
    double s1, c1, c3;
    sincos(two_pi * u1, &s1, &c1);
    c3 = cos(two_pi * u3);
    return make_double3(a * c1, a * s1, b * c3);
}

// ---------------------------------------------------------------------------
// electron heat capacity C = 3 tanh(bT), energy E = 3/b log cosh(bT) and its
// inverse T = acosh(exp(bE/3))/b, b = 2e-4/k_B (heatcapacity_CPU.cpp:34-57).
// For bT <= kPolyXMax (T <= 6250 K) they are evaluated as x^2 Q(x^2),
// x R(x^2) and sqrt(2y) G(y) with the fitted polynomials of heatcap_poly.h
// (about 1 ulp, a few dozen FMAs instead of exp, log, expm1 and log1p);
// outside that range, and for NaN, through the library functions.
template <int N>
__device__ __forceinline__ double horner(const double (&c)[N], double t) {
    double a = c[N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; --k) a = fma(a, t, c[k]);
    return a;
}
constexpr double kCeB = 2e-4 / kBoltz;

// tanh(x), x = bT, for the heat capacity 3 tanh(x)
__device__ __forceinline__ double tanh_bt(double x) {
    if (fabs(x) <= kPolyXMax) return x * horner(kTanhR, x * x);
    const double v = expm1(-2.0 * x);
    return -v / (2.0 + v);
}
// E(T): cosh through one exp and its reciprocal outside the fitted range
__device__ __forceinline__ double te_to_ee(double te) {
    const double x = kCeB * te;
    if (fabs(x) <= kPolyXMax) return (3.0 / kCeB) * (x * x) * horner(kLogCoshQ, x * x);
    const double e = exp(x);
    return 3.0 / kCeB * log(0.5 * (e + 1.0 / e));
}
// T(E); outside the fitted range acosh(e^y) = y + log1p(sqrt(-expm1(-2y)))
__device__ __forceinline__ double ee_to_te(double ee) {
    const double y = kCeB / 3.0 * ee;
    if (y <= kPolyYMax) return sqrt(2.0 * y) * horner(kAcoshExpG, y) * (1.0 / kCeB);
    return (y + log1p(sqrt(-expm1(-2.0 * y)))) / kCeB;
}

}  // namespace sld
