/* Stand-in for LAMMPS's NeighRequest (the pre-2022 interface the reference
   styles use): a request is a half list unless the style sets full. */
#ifndef LMP_NEIGH_REQUEST_H
#define LMP_NEIGH_REQUEST_H

namespace LAMMPS_NS {
class NeighRequest {
   public:
    void *requestor = nullptr;
    int requestor_instance = 0;
    int id = 0;
    int pair = 1;
    int half = 1, full = 0;
    int newton = 0;
};
}  // namespace LAMMPS_NS

#endif
