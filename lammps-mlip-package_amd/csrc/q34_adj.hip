// Order-3/4 kernel instantiation (adjoint G = dE/da), unrolled signatures.
#include "q34_impl.h"

namespace mlip {

hipError_t launch_q34_adj(const DevModel& m, const DevSystem& s, hipStream_t st) {
    if (s.nblocks == 0) return hipSuccess;
    hipError_t e = q34::launch<false, false>(m, s, st);
    if (e == hipSuccess && m.q34_dyn_adj) e = launch_q34_dyn(m, s, false, st);
    return e;
}

}  // namespace mlip
