// Host-side potential tables (see sld_tables.cpp).
#pragma once

#include <memory>
#include <vector>

namespace sld {

struct HostTables {
    double rmax = 0, rhomax = 0, in_over_rmax = 0, in_over_rhomax = 0;
    std::vector<double> rtab;    // [kNInput][kRStride]
    std::vector<double> rhotab;  // [kNInput][kRhoStride]
};

void build_tables(double rcut_max, HostTables& t);
// The tables for rcut_max, built once per process and shared by every
// context created with the same cutoff.
std::shared_ptr<const HostTables> shared_tables(double rcut_max);

}  // namespace sld
