/* Stand-in for LAMMPS's Error: a fatal error prints "ERROR: ..." and ends
   the process with status 3 (LAMMPS calls MPI_Abort / exit). */
#ifndef LMP_ERROR_H
#define LMP_ERROR_H

#include <string>

#include "pointers.h"

namespace LAMMPS_NS {
class Error : protected Pointers {
   public:
    explicit Error(LAMMPS *lmp) : Pointers(lmp) {}
    [[noreturn]] void all(const std::string &file, int line, const std::string &str);
    [[noreturn]] void one(const std::string &file, int line, const std::string &str);
};
}  // namespace LAMMPS_NS

#endif
