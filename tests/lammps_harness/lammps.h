/* Stand-in for LAMMPS's top-level class: the subsystems a pair style uses. */
#ifndef LMP_LAMMPS_H
#define LMP_LAMMPS_H

namespace LAMMPS_NS {
class Memory;
class Error;
class Atom;
class Force;
class Neighbor;

class LAMMPS {
   public:
    LAMMPS();
    ~LAMMPS();
    Memory *memory;
    Error *error;
    Atom *atom;
    Force *force;
    Neighbor *neighbor;
};
}  // namespace LAMMPS_NS

#endif
