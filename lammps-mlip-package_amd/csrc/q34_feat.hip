// Order-3/4 kernel instantiation (polynomial features), unrolled signatures.
#include "q34_impl.h"

namespace mlip {

hipError_t launch_q34_feat(const DevModel& m, const DevSystem& s, hipStream_t st) {
    if (s.nblocks == 0) return hipSuccess;
    hipError_t e = q34::launch<true, false>(m, s, st);
    if (e == hipSuccess && m.q34_dyn_feat) e = launch_q34_dyn(m, s, true, st);
    return e;
}

}  // namespace mlip
