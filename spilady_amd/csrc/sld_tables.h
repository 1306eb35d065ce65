// Host-side potential tables (see sld_tables.cpp).
#pragma once

#include <memory>
#include <vector>

namespace sld {

struct HostTables {
    double rmax = 0, rhomax = 0, in_over_rmax = 0, in_over_rhomax = 0;
    std::vector<double> rtab;    // [kNInput][kRStride]
    std::vector<double> prow;    // [kNInput][kPRowStride]: (rtab[n-1], rtab[n]) per column
    std::vector<double> rhotab;  // [kNInput][kRhoStride]
};

// magmom: the J and Landau columns scaled by powers of the electron g factor
// (potential_CPU.cpp:125-160)
void build_tables(double rcut_max, bool magmom, HostTables& t);

// pieces of the DD05 pair function shared with the parametric model
double zbl_pair(double r);                 // DD05.cpp:45-62
struct ZblEnds {
    double y1, Dy1, D2y1;                  // ZBL value, slope, curvature at r = 1.8
};
struct Bridge {
    double a[6];                           // quintic coefficients, a[k] r^k
};
ZblEnds zbl_ends();
Bridge bridge_from_ends(const ZblEnds& e, double y2, double Dy2, double D2y2);
// The tables for rcut_max (and magmom), built once per process and shared by
// every context created with the same cutoff.
std::shared_ptr<const HostTables> shared_tables(double rcut_max, bool magmom);

}  // namespace sld
