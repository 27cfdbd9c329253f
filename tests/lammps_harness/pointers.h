/* Stand-in for LAMMPS's Pointers: references to the LAMMPS object's
   subsystem pointers, as every LAMMPS class holds them. */
#ifndef LMP_POINTERS_H
#define LMP_POINTERS_H

#include "lammps.h"
#include "lmptype.h"

namespace LAMMPS_NS {
class Pointers {
   public:
    explicit Pointers(LAMMPS *ptr)
        : lmp(ptr), memory(ptr->memory), error(ptr->error), atom(ptr->atom), force(ptr->force),
          neighbor(ptr->neighbor) {}
    virtual ~Pointers() = default;

   protected:
    LAMMPS *lmp;
    Memory *&memory;
    Error *&error;
    Atom *&atom;
    Force *&force;
    Neighbor *&neighbor;
};
}  // namespace LAMMPS_NS

#endif
