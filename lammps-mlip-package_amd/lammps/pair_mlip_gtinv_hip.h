/* LAMMPS pair style `mlip_gtinv/hip`: the reference `mlip_gtinv` contract
   (src/USER-MLIP/pair_mlip_gtinv.h) served by the HIP engine through the C ABI
   of include/mlip_hip.h.  Built only inside a LAMMPS source tree. */
#ifdef PAIR_CLASS
// clang-format off
PairStyle(mlip_gtinv/hip, PairMLIPGtinvHIP)
// clang-format on
#else

#ifndef LMP_PAIR_MLIP_GTINV_HIP_H
#define LMP_PAIR_MLIP_GTINV_HIP_H

#include "mlip_hip.h"
#include "pair.h"

namespace LAMMPS_NS {

class PairMLIPGtinvHIP : public Pair {
   public:
    PairMLIPGtinvHIP(class LAMMPS *);
    ~PairMLIPGtinvHIP() override;
    void compute(int, int) override;
    void settings(int, char **) override;
    void coeff(int, char **) override;
    double init_one(int, int) override;
    void init_style() override;

   protected:
    mlip_ctx *ctx = nullptr;
    int full_list = 0;
    void allocate();
};

}  // namespace LAMMPS_NS

#endif
#endif
