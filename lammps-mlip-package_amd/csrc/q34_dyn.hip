// Order-3/4 kernel instantiation: general-loop visits (signatures outside
// the unrolled set) for both passes.
#include "q34_impl.h"

namespace mlip {

hipError_t launch_q34_dyn(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st) {
    if (s.nblocks == 0) return hipSuccess;
    return feat ? q34::launch<true, true>(m, s, st) : q34::launch<false, true>(m, s, st);
}

}  // namespace mlip
