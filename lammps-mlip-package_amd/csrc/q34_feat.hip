// Order-3/4 kernel instantiation (polynomial features), unrolled signatures.
#include "q34_impl.h"

namespace mlip {

void launch_q34_feat(const DevModel& m, const DevSystem& s, hipStream_t st) {
    if (s.nblocks == 0) return;
    q34::launch<true, false>(m, s, st);
    if (m.q34_dyn_feat) launch_q34_dyn(m, s, true, st);
}

}  // namespace mlip
