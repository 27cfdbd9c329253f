/* Stand-in for the part of LAMMPS's lmptype.h that a pair style sees: the
   default (smallbig) integer sizes and FLERR.  Test harness only: it lets
   lammps-mlip-package_amd/lammps/pair_mlip_gtinv_hip.cpp be compiled and
   driven without a LAMMPS tree. */
#ifndef LMP_LMPTYPE_H
#define LMP_LMPTYPE_H

#include <cstdint>

namespace LAMMPS_NS {
typedef int tagint;      // 32-bit atom IDs unless LAMMPS is built with -DLAMMPS_BIGBIG
typedef int64_t bigint;
}  // namespace LAMMPS_NS

#define FLERR __FILE__, __LINE__

#endif
