/* Stand-in for LAMMPS's Neighbor: the request table pair styles fill in
   init_style (neighbor->request(this, instance_me), then requests[i]). */
#ifndef LMP_NEIGHBOR_H
#define LMP_NEIGHBOR_H

#include "pointers.h"

namespace LAMMPS_NS {
class NeighRequest;
class Neighbor : protected Pointers {
   public:
    explicit Neighbor(LAMMPS *lmp) : Pointers(lmp) {}
    ~Neighbor() override;
    int nrequest = 0;
    NeighRequest *requests[8] = {};
    int request(void *requestor, int instance = 0);
};
}  // namespace LAMMPS_NS

#endif
