/* Stand-in for LAMMPS's Atom: per-atom arrays of owned atoms followed by
   ghosts (nlocal + nghost entries), 1-based types, masses per type. */
#ifndef LMP_ATOM_H
#define LMP_ATOM_H

#include "pointers.h"

namespace LAMMPS_NS {
class Atom : protected Pointers {
   public:
    explicit Atom(LAMMPS *lmp) : Pointers(lmp) {}
    ~Atom() override;
    int nlocal = 0, nghost = 0, nmax = 0, ntypes = 0;
    int tag_enable = 1;
    double **x = nullptr, **f = nullptr;
    int *type = nullptr;
    tagint *tag = nullptr;
    double *mass = nullptr;      // [ntypes + 1]
    int *mass_setflag = nullptr;  // [ntypes + 1]
    void grow(int n, int ntype);
    void set_mass(const char *file, int line, int itype, double value);
};
}  // namespace LAMMPS_NS

#endif
