// Synthetic batched/served structures (see workload.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace mlip {

// log-uniform size in [nmin, nmax] of structure `index`
int random_structure_size(uint64_t seed, int index, int nmin, int nmax);
// random triclinic cell (rows a, b, c) at `density` with positions under an
// rmin rejection and 1-based types; false if the rejection cannot finish
bool random_structure(uint64_t seed, int index, int natoms, double density, double rmin, double cell[9], double* x,
                      int* type);

}  // namespace mlip
