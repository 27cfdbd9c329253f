/* LAMMPS pair styles `mlip_gtinv/hip` and `mlip_pair/hip`: the reference
   `mlip_gtinv` and `mlip_pair` contracts (src/USER-MLIP/pair_mlip_gtinv.h,
   pair_mlip_pair.h) served by the HIP engine through the C ABI of
   include/mlip_hip.h; the engine reads the style from the potential file's
   des_type line.  Built only inside a LAMMPS source tree. */
#ifdef PAIR_CLASS
// clang-format off
PairStyle(mlip_gtinv/hip, PairMLIPGtinvHIP)
PairStyle(mlip_pair/hip, PairMLIPGtinvHIP)
// clang-format on
#else

#ifndef LMP_PAIR_MLIP_GTINV_HIP_H
#define LMP_PAIR_MLIP_GTINV_HIP_H

#include <cstdint>
#include <vector>

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
    std::vector<int64_t> tag64;  // atom->tag widened to the ABI's int64
    void allocate();
};

}  // namespace LAMMPS_NS

#endif
#endif
