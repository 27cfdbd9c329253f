// Order-3/4 kernel instantiation: adjoint G = dE/da.
#include <cstdio>

#include "q34_impl.h"

namespace mlip {

void launch_q34_adj(const DevModel& m, const DevSystem& s, hipStream_t st) {
    if (s.nblocks == 0) return;
    const size_t lds = sizeof(double) * m.n_type * m.n_lmh3 * 2 * kBlock;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(q34::k_q34<false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(q34::k_q34<false>, dim3(s.nblocks, m.n_fn), dim3(q34::kQ34Threads), lds, st, m, s);
        e = hipGetLastError();
    }
    if (e != hipSuccess) std::fprintf(stderr, "mlip_hip: k_q34 launch failed: %s\n", hipGetErrorString(e));
}

}  // namespace mlip
