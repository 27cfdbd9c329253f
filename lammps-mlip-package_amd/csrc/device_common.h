// Device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "device_api.h"

namespace mlip {
namespace dev {

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


}  // namespace dev
}  // namespace mlip
