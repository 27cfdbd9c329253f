/* Stand-in for LAMMPS's Memory: 1-D arrays and 2-D arrays as one
   contiguous block behind row pointers (so &a[0][0] walks all rows). */
#ifndef LMP_MEMORY_H
#define LMP_MEMORY_H

#include <cstdlib>

#include "pointers.h"

namespace LAMMPS_NS {
class Memory : protected Pointers {
   public:
    explicit Memory(LAMMPS *lmp) : Pointers(lmp) {}
    template <typename T>
    T *create(T *&array, int n, const char *) {
        array = static_cast<T *>(calloc(n > 0 ? n : 1, sizeof(T)));
        return array;
    }
    template <typename T>
    T **create(T **&array, int n1, int n2, const char *) {
        T *data = static_cast<T *>(calloc((n1 * n2) > 0 ? n1 * n2 : 1, sizeof(T)));
        array = static_cast<T **>(calloc(n1 > 0 ? n1 : 1, sizeof(T *)));
        for (int i = 0; i < n1; ++i) array[i] = data + static_cast<size_t>(i) * n2;
        return array;
    }
    template <typename T>
    void destroy(T *&array) {
        free(array);
        array = nullptr;
    }
    template <typename T>
    void destroy(T **&array) {
        if (array) free(array[0]);
        free(array);
        array = nullptr;
    }
};
}  // namespace LAMMPS_NS

#endif
