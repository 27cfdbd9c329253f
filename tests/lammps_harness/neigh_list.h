/* Stand-in for LAMMPS's NeighList: ilist of owned atoms, per-atom counts
   and per-atom pointers into the neighbor pages. */
#ifndef LMP_NEIGH_LIST_H
#define LMP_NEIGH_LIST_H

namespace LAMMPS_NS {
class NeighList {
   public:
    int inum = 0, gnum = 0;
    int *ilist = nullptr;
    int *numneigh = nullptr;
    int **firstneigh = nullptr;
};
}  // namespace LAMMPS_NS

#endif
