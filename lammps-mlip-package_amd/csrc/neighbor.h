// Host-side periodic neighbor list builder (full list, owned atoms, shifts).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace mlip {

struct NeighborList {
    std::vector<int64_t> off;     // n+1
    std::vector<int32_t> idx;     // neighbor owned-atom index
    std::vector<double> shift;    // 3 per pair: del = x_i - x_j + shift
    std::vector<double> wrap_shift;
};

// Bin grid of a periodic cell for cutoff rc: inverse cell, bins per lattice
// direction (each at least rc/2 wide across its faces) and the stencil
// reach in bins.  Shared by the host builder and the device builder, so both
// bin and wrap identically.
constexpr int64_t kMaxCellBins = int64_t(1) << 28;
struct CellBins {
    double g[9];     // inverse of the column matrix [a b c]
    double vol;
    int nb[3];
    int reach[3];
};
bool cell_bins(const double cell[9], double rc, CellBins& out, std::string& err);

// cell: rows a, b, c.  Pairs with 0 < r < rc, self-images included; each
// atom's neighbors in a fixed order whatever nthreads (atoms are split
// into contiguous ranges across threads for n >= 512).
bool build_full_list(const double cell[9], int n, const double* x, double rc, NeighborList& out, std::string& err,
                     int nthreads = 1);

}  // namespace mlip
