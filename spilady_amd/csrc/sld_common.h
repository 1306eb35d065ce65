// Constants and capacities shared by host and device code.
#pragma once

namespace sld {

// spilady.h:53-62
constexpr double kBoltz = 8.617343e-5;       // eV/K
constexpr double kHbar = 6.58211899e-16;     // eV s
constexpr double kElG = 2.0023193043622e0;   // electron g factor
constexpr double kEvToJ = 1.602176487e-19;
constexpr double kPi = 3.141592653589793238462e0;  // Pi_num (spilady.h:61)

// potential tables (global.h:66): 10^6 points per table
constexpr int kNInput = 1000000;
constexpr int kRStride = 8;     // sf dsf pr dpr J dJ pad pad
constexpr int kRhoStride = 16;  // F dF A B C D dA dB dC dD pad...

enum RCol { R_SF = 0, R_DSF, R_PR, R_DPR, R_J, R_DJ };
// Device r table, one 128-byte row per lookup index n holding the pair
// (t[n-1], t[n]) that the interpolation at n reads, for each column: a pair
// term's lookup touches one cache line instead of two 64-byte rows.
constexpr int kPRowStride = 16;
enum PCol { P_DSF = 0, P_DPR = 2, P_PR = 4, P_J = 6, P_DJ = 8, P_SF = 10 };
enum RhoCol { H_F = 0, H_DF, H_A, H_B, H_C, H_D, H_DA, H_DB, H_DC, H_DD };

// Fixed capacities; exceeding one sets an error flag that the host turns
// into SLD_ERR_CAPACITY.
constexpr int kCellCap = 32;    // atoms per link cell
constexpr int kMaxNb = 48;      // neighbours within rcut_max per atom
constexpr int kMaxMag = 32;     // neighbours within rcut_mag per atom

// error flag bits
constexpr int kErrCellOverflow = 1;
constexpr int kErrNbOverflow = 2;
constexpr int kErrFarMove = 4;
constexpr int kErrNonFinite = 8;

}  // namespace sld
