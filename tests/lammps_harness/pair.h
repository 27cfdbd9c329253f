/* Stand-in for LAMMPS's Pair base class: the flags, accumulators and
   per-atom arrays a pair style uses, and ev_setup with the semantics of
   LAMMPS's Pair::ev_setup for the flag bits its integrators pass. */
#ifndef LMP_PAIR_H
#define LMP_PAIR_H

#include "pointers.h"

namespace LAMMPS_NS {
// eflag / vflag bits as Integrate::ev_set passes them to Pair::compute
enum { ENERGY_NONE = 0, ENERGY_GLOBAL = 1, ENERGY_ATOM = 2 };
enum { VIRIAL_NONE = 0, VIRIAL_PAIR = 1, VIRIAL_FDOTR = 2, VIRIAL_ATOM = 4 };

class NeighList;

class Pair : protected Pointers {
   public:
    explicit Pair(LAMMPS *lmp) : Pointers(lmp) {}
    ~Pair() override;

    double eng_vdwl = 0.0, eng_coul = 0.0;
    double virial[6] = {0, 0, 0, 0, 0, 0};
    double *eatom = nullptr;
    double **vatom = nullptr;

    int single_enable = 1, restartinfo = 1, one_coeff = 0, manybody_flag = 0;
    int no_virial_fdotr_compute = 0;
    int copymode = 0, allocated = 0, instance_me = 0;
    int **setflag = nullptr;
    double **cutsq = nullptr;

    int evflag = 0, eflag_either = 0, eflag_global = 0, eflag_atom = 0;
    int vflag_either = 0, vflag_global = 0, vflag_atom = 0, vflag_fdotr = 0;

    NeighList *list = nullptr;

    virtual void compute(int, int) = 0;
    virtual void settings(int, char **) = 0;
    virtual void coeff(int, char **) = 0;
    virtual double init_one(int, int) { return 0.0; }
    virtual void init_style();
    virtual void init_list(int, NeighList *ptr) { list = ptr; }

   protected:
    int maxeatom = 0, maxvatom = 0;
    void ev_init(int eflag, int vflag, int alloc = 1) {
        if (eflag || vflag)
            ev_setup(eflag, vflag, alloc);
        else
            ev_unset();
    }
    void ev_setup(int eflag, int vflag, int alloc = 1);
    void ev_unset();
    void virial_fdotr_compute();
};
}  // namespace LAMMPS_NS

#endif
