// Device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "device_api.h"

namespace mlip {
namespace dev {

// Read-only tables (coefficients, visit descriptors) are addressed through
// the constant address space: wave-uniform reads of it compile to scalar
// loads through the scalar cache instead of per-lane vector loads.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
#else
template <typename T>
using cptr = const T*;  // host pass: kernels are not compiled for the host
#endif

// Pointer asserted wave-uniform (both 32-bit halves through readfirstlane),
// so that loads through it stay scalar even across non-inlined calls.
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v >> 32));
    return (cptr<T>)(reinterpret_cast<const T*>((static_cast<unsigned long long>(hi) << 32) | lo));
}

__device__ __forceinline__ int lmh(int l, int m) { return l * (l + 1) / 2 + l + m; }

// XCD-aware mapping of a 1-D grid onto slots: consecutive workgroups of one
// XCD (workgroups are dealt round-robin over the 8 XCDs) walk consecutive
// lanes of the same slot block, so the block-interleaved a/G rows they touch
// share that XCD's L2 lines.
__device__ __forceinline__ int xcd_slot(int wg, int nblocks) {
    const int xcd = wg & 7;
    const int k = wg >> 3;
    const int blk = xcd + 8 * (k / kBlock);
    if (blk >= nblocks) return -1;
    return blk * kBlock + (k % kBlock);
}

__device__ __forceinline__ size_t a_index(const DevModel& M, int blk, int n, int s, int lm, int c, int lane) {
    return (((((size_t)blk * M.n_fn + n) * M.n_type + s) * M.n_lmh + lm) * 2 + c) * kBlock + lane;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Column of polynomial feature (n, rank) in the compact row of a centre of
// species t (the rank lies in the species' range prange[t]).
__device__ __forceinline__ int pcol(const DevModel& M, int t, int n, int rank) {
    return n * (M.prange[t][1] - M.prange[t][0]) + rank - M.prange[t][0];
}


}  // namespace dev
}  // namespace mlip
