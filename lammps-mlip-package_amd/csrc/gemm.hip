// Quadratic form of the polynomial model on the matrix cores (FP64 MFMA).
//
// Per centre species t, for the species' slot rows (blocks are grouped by
// species, and a row holds only the features the species sees):
//   Hq[slot][a] = sum_b D[slot][b] W_t[b][a]
// with W_t the symmetric quadratic form of the species' features
// (reference: the comb2 part of compute_polynomial_model_uniq_products and
// compute_anlm_polynomial_model_correction, pair_mlip_gtinv.cpp:321-343,
// 481-521, and compute_polynomial_adjoints, mlipkk.cpp:668-731).
//
// Workgroup tile 64 rows x 128 columns, K staged through LDS 16 at a time
// (the next stage's global loads in flight while the current one computes);
// four waves of 32 x 64 each, i.e. 2 x 4 blocks of v_mfma_f64_16x16x4f64.
// Lane layout of that instruction (measured on gfx950): A[i][k] from lane
// 16 k + i, B[k][j] from lane 16 k + j, result register r of lane l is
// D[l / 16 + 4 r][l % 16].
#include <hip/hip_runtime.h>

#include "device_api.h"
#include "device_common.h"

namespace mlip {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kGM = 64;        // rows per workgroup
constexpr int kGN = 128;       // columns per workgroup
constexpr int kGK = 16;        // K per LDS stage
constexpr int kGThreads = 256;
// LDS row padding (doubles): rows of the K slabs 80 and 144 doubles apart
// (measured 2.48 -> 2.35 ms per evaluation at 256k against a pad of 1)
constexpr int kGPad = 16;

// EPI: the polynomial step of a model without cubic terms in the epilogue,
// C = H = lin + A B and epart[row * ldep + nt] = sum over the tile's columns
// of A[row][col] (lin + (A B) / 2) (A holds the features d, K = N), each row's
// partial in a fixed order: four columns per lane, a butterfly over the 16
// lanes of the row, then the two column waves.  Held to 4 waves per SIMD
// (128 registers with the accumulators): unconstrained, the compiler issues
// every epilogue load at once (108 VGPRs), which halves the occupancy of
// the product loop (2.35 -> 2.97 ms at 256k).
template <bool EPI>
__global__ __launch_bounds__(kGThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_gemm_rows(const double* __restrict__ A, int lda,
                                                         const double* __restrict__ B, int ldb,
                                                         double* __restrict__ C, int ldc, int K, int N,
                                                         int mtiles, int ntiles, const double* __restrict__ lin,
                                                         double* __restrict__ epart, int ldep) {
    __shared__ double As[kGK][kGM + kGPad];  // [k][row]
    __shared__ double Bs[kGK][kGN + kGPad];  // [k][column]
    // XCD-aware order: workgroups are dealt round-robin over the 8 XCDs, and
    // consecutive workgroups of one XCD take the column tiles of one row
    // tile, so its A rows are fetched into that XCD's L2 once
    const int g = blockIdx.x;
    const int xcd = g & 7, k8 = g >> 3;
    const int mt = xcd + 8 * (k8 / ntiles), nt = k8 % ntiles;
    if (mt >= mtiles) return;
    const int m0 = mt * kGM, n0 = nt * kGN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    // global -> register staging roles
    const int arow = tid >> 2, ak = (tid & 3) * 4;   // A: 64 rows x 16 k, 4 per thread
    // B: 16 k x 128 columns, 8 per thread, 16 apart (a row's 16 lanes store
    // consecutive doubles: no LDS bank conflicts; global reads stay 128-byte runs)
    const int bk = tid >> 4, bc = tid & 15;
    const double* arow_p = A + (size_t)(m0 + arow) * lda;
    double ra[4], rb[8];
    auto load = [&](int kb) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = kb + ak + q;
            ra[q] = k < K ? arow_p[k] : 0.0;
        }
        const int k = kb + bk;
        const double* brow = B + (size_t)k * ldb;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = n0 + bc + 16 * q;
            rb[q] = (k < K && c < N) ? brow[c] : 0.0;
        }
    };
    dbl4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
    const int nk = (K + kGK - 1) / kGK;
    load(0);
    for (int kt = 0; kt < nk; ++kt) {
        if (kt) __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) As[ak + q][arow] = ra[q];
#pragma unroll
        for (int q = 0; q < 8; ++q) Bs[bk][bc + 16 * q] = rb[q];
        __syncthreads();
        if (kt + 1 < nk) load((kt + 1) * kGK);
#pragma unroll
        for (int kk = 0; kk < kGK / 4; ++kk) {
            const int ka = kk * 4 + (lane >> 4);
            double a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[ka][wr * 32 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[ka][wc * 64 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    if constexpr (EPI) {
        __shared__ double esum[2][kGM];  // [column wave][row of the tile]
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rt = wr * 32 + i * 16 + (lane >> 4) + 4 * r;
                const size_t row = static_cast<size_t>(m0 + rt);
                double e = 0.0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int col = n0 + wc * 64 + j * 16 + (lane & 15);
                    if (col < N) {
                        const double la = lin[col], hq = acc[i][j][r];
                        C[row * ldc + col] = la + hq;
                        e = fma(A[row * lda + col], la + 0.5 * hq, e);
                    }
                }
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) e += __shfl_xor(e, o);
                if ((lane & 15) == 0) esum[wc][rt] = e;
            }
        __syncthreads();
        if (tid < kGM) epart[static_cast<size_t>(m0 + tid) * ldep + nt] = esum[0][tid] + esum[1][tid];
        return;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wc * 64 + j * 16 + (lane & 15);
            if (col >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wr * 32 + i * 16 + (lane >> 4) + 4 * r;
                C[(size_t)row * ldc + col] = acc[i][j][r];
            }
        }
}

// Ep[slot] = the slot's partial energies of the fused epilogue, in column
// tile order (no tile for a species without polynomial features: 0)
__global__ __launch_bounds__(256) void k_ep_sum(DevModel M, DevSystem S) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= S.nblocks * kBlock) return;
    const int t = S.block_type[slot / kBlock];
    const int nt = (M.wdim[t] + kGN - 1) / kGN;
    const double* ep = S.Hq + static_cast<size_t>(slot) * M.n_pr;
    double e = 0.0;
    for (int k = 0; k < nt; ++k) e += ep[k];
    S.Ep[slot] = e;
}

}  // namespace

hipError_t launch_poly_gemm(const DevModel& m, const DevSystem& s, const int* nblk_type, hipStream_t st) {
    int b0 = 0;
    for (int t = 0; t < m.n_type && t < 2; ++t) {
        const int nb = nblk_type[t], w = m.wdim[t];
        const size_t r0 = static_cast<size_t>(b0) * kBlock;
        b0 += nb;
        if (nb == 0 || w == 0) continue;
        const int mtiles = nb * kBlock / kGM, ntiles = (w + kGN - 1) / kGN;
        const int grid = 8 * ntiles * ((mtiles + 7) / 8);
        if (m.poly_fused)
            hipLaunchKernelGGL(k_gemm_rows<true>, dim3(grid), dim3(kGThreads), 0, st, s.D + r0 * m.n_pr, m.n_pr,
                               m.wmat + m.wofs[t], w, s.H + r0 * m.n_pr, m.n_pr, w, w, mtiles, ntiles,
                               m.lin_p + static_cast<size_t>(t) * m.n_pr, s.Hq + r0 * m.n_pr, m.n_pr);
        else
            hipLaunchKernelGGL(k_gemm_rows<false>, dim3(grid), dim3(kGThreads), 0, st, s.D + r0 * m.n_pr, m.n_pr,
                               m.wmat + m.wofs[t], w, s.Hq + r0 * m.n_pr, m.n_pr, w, w, mtiles, ntiles,
                               nullptr, nullptr, 0);
        const hipError_t e = launched();
        if (e != hipSuccess) return e;
    }
    return b0 == s.nblocks ? hipSuccess : hipErrorInvalidValue;
}

hipError_t launch_poly_sum(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ep_sum, dim3((nslot + 255) / 256), dim3(256), 0, st, m, s);
    return launched();
}

}  // namespace mlip
