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

// cell: rows a, b, c.  Pairs with 0 < r < rc, self-images included; each
// atom's neighbors in a fixed order whatever nthreads (atoms are split
// into contiguous ranges across threads for n >= 512).
bool build_full_list(const double cell[9], int n, const double* x, double rc, NeighborList& out, std::string& err,
                     int nthreads = 1);

}  // namespace mlip
