// HIP kernels of the gtinv energy/force path for CDNA4 (gfx950).
//
// Pipeline for one evaluation (reference: PairMLIPGtinv::compute,
// src/USER-MLIP/pair_mlip_gtinv.cpp:63-125, and docs/formulation.md):
//   k_anlm   a_{i,s,n,lm} = sum_j g_n(r_ij) Y_lm(r_ij)          (:349-447, pair_kernels.hip)
//   k_q12    order-1/2 invariants: polynomial features or adjoint G
//   k_q34    order-3/4 invariants: polynomial features or adjoint G,
//            a_nlm tile of 64 same-species centres resident in LDS
//   k_gemm_rows  Hq = D W (quadratic form of the polynomial model, FP64 MFMA; gemm.hip),
//                or, without cubic terms, H and the polynomial energy in its epilogue
//   k_poly   E_i polynomial part and H = dE_i/dd                 (:321-343,481-521)
//   k_force  per-pair dE_i/d(del) from G and dphi/d(del)          (:161-256, pair_kernels.hip)
//   k_gather atomics-free per-position force sums, deterministic
//   k_energy per-centre energies and virial, fixed-order reductions
//
// Conventions shared with the reference CPU style: del = x_centre - x_neighbor,
// a_nlm uses Y_lm(del) without conjugation, pairs with r < cutoff (strict).
// Only m <= 0 entries of a and G are stored (index l(l+1)/2 + l + m); the
// m > 0 entries follow from a_{l,m} = (-1)^m conj(a_{l,-m}).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "device_api.h"
#include "device_common.h"

namespace mlip {

namespace {

using namespace dev;


// ---------------------------------------------------------------- order 1/2
constexpr int kQ12Pitch = kBlock + 1;           // LDS tile pitch (doubles) of the staged feature pass
constexpr size_t kQ12MaxStage = 32 * 1024;      // largest staged tile (bytes)
// Wave = 64 slots of one block at one radial index n; a_{s,L,m} for the
// current L is held in registers, G_{s,L,m} accumulated in registers.
template <int L, bool FEAT>
__device__ void q12_level(const DevModel& M, const DevSystem& S, int blk, int n, int lane, int t,
                          double& eacc, double* tile) {
    double ar[2][L + 1], ai[2][L + 1], gr[2][L + 1], gi[2][L + 1];
    const int ns = M.n_type;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int k = 0; k <= L; ++k) {
            gr[s][k] = gi[s][k] = 0.0;
            if (s < ns) {
                ar[s][k] = S.A[a_index(M, blk, n, s, lmh(L, -k), 0, lane)];
                ai[s][k] = S.A[a_index(M, blk, n, s, lmh(L, -k), 1, lane)];
            } else {
                ar[s][k] = ai[s][k] = 0.0;
            }
        }
    const int slot = blk * kBlock + lane;
    const int vb = M.q12_off[t * (M.maxl + 2) + L], ve = M.q12_off[t * (M.maxl + 2) + L + 1];
    for (int vi = vb; vi < ve; ++vi) {
        const VisitQ12 v = M.q12[vi];
        const int pslot = uniform(v.pslot);
        if (FEAT && pslot < 0) continue;
        const int s0 = uniform(v.s0), s1 = uniform(v.s1);
        const double* cp = M.coef + uniform(v.coef);
        double H = 0.0;
        if (!FEAT) H = pslot >= 0 ? S.H[(size_t)slot * M.n_pr + pcol(M, t, n, pslot)] : M.weight[n * M.n_gtinv + v.itc];
        const double f0 = s0 == 0 ? 1.0 : 0.0, f1 = 1.0 - f0;
        const double e0 = s1 == 0 ? 1.0 : 0.0, e1 = 1.0 - e0;
        double d = 0.0;
        if (uniform(v.order) == 1) {
            // d = C a_{s,0,0} (real); dd/da_{s,0,0} = C
            const double c = cp[0];
            d = c * (f0 * ar[0][0] + f1 * ar[1][0]);
            if (!FEAT) {
                gr[0][0] = fma(f0 * H, c, gr[0][0]);
                gr[1][0] = fma(f1 * H, c, gr[1][0]);
            }
        } else {
#pragma unroll
            for (int k = 0; k <= L; ++k) {
                const double sg = (k & 1) ? -1.0 : 1.0;
                // factor values: x = a_{s0,L,-k}, y = a_{s1,L,-k}
                const double xr = f0 * ar[0][k] + f1 * ar[1][k], xi = f0 * ai[0][k] + f1 * ai[1][k];
                const double yr = e0 * ar[0][k] + e1 * ar[1][k], yi = e0 * ai[0][k] + e1 * ai[1][k];
                const double cm = cp[L - k], cpk = cp[L + k];  // C(m0=-k), C(m0=+k)
                // pair-summed real part: Re(a_{s0,-k} a_{s1,+k}) with a_{+k} = sg conj(a_{-k})
                const double re = sg * (xr * yr + xi * yi);
                d = fma(k == 0 ? cm : cm + cpk, re, d);
                if (!FEAT) {
                    // G_{s0,L,-k} += H C(-k) a_{s1,L,+k};  G_{s1,L,-k} += H C(+k) a_{s0,L,+k}
                    const double h0 = H * cm * sg, h1 = H * cpk * sg;
                    gr[0][k] = fma(f0 * h0, yr, gr[0][k]);
                    gi[0][k] = fma(-f0 * h0, yi, gi[0][k]);
                    gr[1][k] = fma(f1 * h0, yr, gr[1][k]);
                    gi[1][k] = fma(-f1 * h0, yi, gi[1][k]);
                    gr[0][k] = fma(e0 * h1, xr, gr[0][k]);
                    gi[0][k] = fma(-e0 * h1, xi, gi[0][k]);
                    gr[1][k] = fma(e1 * h1, xr, gr[1][k]);
                    gi[1][k] = fma(-e1 * h1, xi, gi[1][k]);
                }
            }
        }
        if (FEAT) {
            if (tile)
                tile[(pslot - M.prange[t][0]) * kQ12Pitch + lane] = d;
            else
                S.D[(size_t)slot * M.n_pr + pcol(M, t, n, pslot)] = d;
        } else if (pslot < 0) {
            eacc = fma(M.weight[n * M.n_gtinv + v.itc], d, eacc);
        }
    }
    if (!FEAT) {
#pragma unroll
        for (int s = 0; s < 2; ++s)
            if (s < ns)
#pragma unroll
                for (int k = 0; k <= L; ++k) {
                    S.G[a_index(M, blk, n, s, lmh(L, -k), 0, lane)] = gr[s][k];
                    S.G[a_index(M, blk, n, s, lmh(L, -k), 1, lane)] = gi[s][k];
                }
    }
}

template <int L, int LMAX, bool FEAT>
struct Q12Levels {
    __device__ static void run(const DevModel& M, const DevSystem& S, int blk, int n, int lane, int t,
                               double& e, double* tile) {
        // fused adjoint levels run in the order-3/4 kernel
        if (L <= M.maxl && (FEAT || !q12_fused(M, L))) q12_level<L, FEAT>(M, S, blk, n, lane, t, e, tile);
        Q12Levels<L + 1, LMAX, FEAT>::run(M, S, blk, n, lane, t, e, tile);
    }
};
template <int LMAX, bool FEAT>
struct Q12Levels<LMAX + 1, LMAX, FEAT> {
    __device__ static void run(const DevModel&, const DevSystem&, int, int, int, int, double&, double*) {}
};

// Levels above kMaxLQ12 (order-2 families up to l = 20): the same visits
// with a and G read and updated in HBM term by term instead of held in
// registers (the row of 2 (L+1) complex values per species would not fit).
template <bool FEAT>
__device__ void q12_level_dyn(const DevModel& M, const DevSystem& S, int blk, int n, int lane, int t, int L,
                              double& eacc, double* tile) {
    const int ns = M.n_type;
    const int slot = blk * kBlock + lane;
    if (!FEAT)
        for (int s = 0; s < ns; ++s)
            for (int k = 0; k <= L; ++k) {
                S.G[a_index(M, blk, n, s, lmh(L, -k), 0, lane)] = 0.0;
                S.G[a_index(M, blk, n, s, lmh(L, -k), 1, lane)] = 0.0;
            }
    const int vb = M.q12_off[t * (M.maxl + 2) + L], ve = M.q12_off[t * (M.maxl + 2) + L + 1];
    for (int vi = vb; vi < ve; ++vi) {
        const VisitQ12 v = M.q12[vi];
        const int pslot = uniform(v.pslot);
        if (FEAT && pslot < 0) continue;
        if (uniform(v.order) != 2) continue;  // order-1 families live at l = 0
        const int s0 = uniform(v.s0), s1 = uniform(v.s1);
        const double* cp = M.coef + uniform(v.coef);
        double H = 0.0;
        if (!FEAT) H = pslot >= 0 ? S.H[(size_t)slot * M.n_pr + pcol(M, t, n, pslot)] : M.weight[n * M.n_gtinv + v.itc];
        double d = 0.0;
        for (int k = 0; k <= L; ++k) {
            const double sg = (k & 1) ? -1.0 : 1.0;
            const int lm = lmh(L, -k);
            const double xr = S.A[a_index(M, blk, n, s0, lm, 0, lane)], xi = S.A[a_index(M, blk, n, s0, lm, 1, lane)];
            const double yr = S.A[a_index(M, blk, n, s1, lm, 0, lane)], yi = S.A[a_index(M, blk, n, s1, lm, 1, lane)];
            const double cm = cp[L - k], cpk = cp[L + k];
            d = fma(k == 0 ? cm : cm + cpk, sg * (xr * yr + xi * yi), d);
            if (!FEAT) {
                const double h0 = H * cm * sg, h1 = H * cpk * sg;
                double* g0 = S.G + a_index(M, blk, n, s0, lm, 0, lane);
                double* g1 = S.G + a_index(M, blk, n, s1, lm, 0, lane);
                const size_t im = kBlock;  // imaginary part: next component, same lane
                g0[0] = fma(h0, yr, g0[0]);
                g0[im] = fma(-h0, yi, g0[im]);
                g1[0] = fma(h1, xr, g1[0]);
                g1[im] = fma(-h1, xi, g1[im]);
            }
        }
        if (FEAT) {
            if (tile)
                tile[(pslot - M.prange[t][0]) * kQ12Pitch + lane] = d;
            else
                S.D[(size_t)slot * M.n_pr + pcol(M, t, n, pslot)] = d;
        } else if (pslot < 0) {
            eacc = fma(M.weight[n * M.n_gtinv + v.itc], d, eacc);
        }
    }
}

// The feature pass (stage = 1) collects the block's row segments of radial
// index n in LDS, [rank - lo][lane] with a pitch of kQ12Pitch, zeroed first
// (features no visit writes stay zero), and writes them out row by row in
// whole lines; without staging each lane writes its own row directly.
template <bool FEAT>
__global__ __launch_bounds__(64) void k_q12(DevModel M, DevSystem S, int stage) {
    extern __shared__ double q12_tile[];
    const int blk = blockIdx.x, n = blockIdx.y, lane = threadIdx.x;
    const int t = S.block_type[blk];
    const int w = M.prange[t][1] - M.prange[t][0];
    double* tile = (FEAT && stage) ? q12_tile : nullptr;
    double e = 0.0;
    if (FEAT) {
        if (tile) {
            for (int k = lane; k < w * kQ12Pitch; k += 64) tile[k] = 0.0;
            __syncthreads();
        } else {
            // features of families the centre's species does not see are zero
            // (the reference never accumulates them); the visits below and the
            // order-3/4 feature kernel write every other slot of the row
            double* row = S.D + (size_t)(blk * kBlock + lane) * M.n_pr + pcol(M, t, n, M.prange[t][0]);
            const int z0 = M.zoff[t], z1 = M.zoff[t + 1];
            for (int r = z0; r < z1; ++r) row[M.zlist[r]] = 0.0;
        }
    }
    Q12Levels<0, kMaxLQ12, FEAT>::run(M, S, blk, n, lane, t, e, tile);
    for (int L = kMaxLQ12 + 1; L <= M.maxl; ++L) q12_level_dyn<FEAT>(M, S, blk, n, lane, t, L, e, tile);
    if (!FEAT) S.Eq12[(size_t)(blk * kBlock + lane) * M.n_fn + n] = e;
    if (FEAT && tile && w > 0) {
        __syncthreads();
        // segment of slot j: w doubles from column n * w of its row
        double* base = S.D + (size_t)blk * kBlock * M.n_pr + (size_t)n * w;
        int j = lane / w, r = lane - j * w;
        for (int k = lane; k < kBlock * w; k += 64) {
            base[(size_t)j * M.n_pr + r] = tile[r * kQ12Pitch + j];
            r += 64;
            while (r >= w) {
                r -= w;
                ++j;
            }
        }
    }
}

// ---------------------------------------------------------------- polynomial
// One wave per slot: lanes stride the polynomial features.
//   H_a = lin_a + (W d)_a + sum_{cubic} w d_b d_c
//   E_P = sum_a d_a (lin_a + Hq_a / 2 + H3_a / 3)   (Euler, homogeneous parts)
__global__ __launch_bounds__(256) void k_poly(DevModel M, DevSystem S) {
    const int slot = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    __shared__ double red[4][64];
    double e = 0.0;
    if (slot < S.nblocks * kBlock) {
        // the compact row of the slot's species (features it does not see
        // are zero and have no column)
        const int t = S.block_type[slot / kBlock];
        const double* d = S.D + (size_t)slot * M.n_pr;
        const double* hq = S.Hq + (size_t)slot * M.n_pr;
        double* h = S.H + (size_t)slot * M.n_pr;
        const int w = M.n_fn * (M.prange[t][1] - M.prange[t][0]);
        const int32_t* c3o = M.c3_off + t * (M.n_pr + 1);
        const double* lin = M.lin_p + (size_t)t * M.n_pr;
        for (int a = lane; a < w; a += 64) {
            double h3 = 0.0;
            for (int k = c3o[a]; k < c3o[a + 1]; ++k)
                h3 = fma(M.c3_w[k], d[M.c3_bc[2 * k]] * d[M.c3_bc[2 * k + 1]], h3);
            const double la = lin[a];
            h[a] = la + hq[a] + h3;
            e = fma(d[a], la + 0.5 * hq[a] + h3 / 3.0, e);
        }
    }
    red[threadIdx.x >> 6][lane] = e;
    __syncthreads();
    if (lane == 0 && slot < S.nblocks * kBlock) {
        double s = 0.0;
        for (int k = 0; k < 64; ++k) s += red[threadIdx.x >> 6][k];
        S.Ep[slot] = s;
    }
}

// Split polynomial adjoint for the per-atom energy (one wave per slot):
//   He_a = lin_a / q_a + sum_b W_ab d_b / (q_a + q_b) + sum_cubic w d_b d_c / (q_a + q_b + q_c)
// with q the degree (family order) of each feature: the reference divides
// every term of dE/da by the term's total degree in a
// (pair_mlip_gtinv.cpp:308-340, mlip_polynomial_gtinv.cpp:147-166).
__global__ __launch_bounds__(256) void k_poly_split(DevModel M, DevSystem S) {
    const int slot = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (slot >= S.nblocks * kBlock) return;
    const int t = S.block_type[slot / kBlock];
    const double* d = S.D + (size_t)slot * M.n_pr;
    double* he = S.Hq + (size_t)slot * M.n_pr;
    const int w = M.n_fn * (M.prange[t][1] - M.prange[t][0]);
    const int32_t* deg = M.pdeg + (size_t)t * M.n_pr;
    const double* W = M.wmat + M.wofs[t];
    const int32_t* c3o = M.c3_off + t * (M.n_pr + 1);
    const double* lin = M.lin_p + (size_t)t * M.n_pr;
    for (int a = lane; a < w; a += 64) {
        const int qa = deg[a];
        double h = lin[a] / qa;
        // W is symmetric: W[b][a] for fixed b is contiguous across lanes
        for (int b = 0; b < w; ++b) h = fma(W[(size_t)b * w + a], d[b] / (qa + deg[b]), h);
        for (int k = c3o[a]; k < c3o[a + 1]; ++k) {
            const int b = M.c3_bc[2 * k], c = M.c3_bc[2 * k + 1];
            h = fma(M.c3_w[k], d[b] * d[c] / (qa + deg[b] + deg[c]), h);
        }
        he[a] = h;
    }
}

// Per position k: half the energy share and virial of every pair in its
// gather list (centre or neighbor side), in list order.
__global__ __launch_bounds__(256) void k_atom_ev(DevSystem S, const double* __restrict__ epv, double* evat) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S.nall) return;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int e = S.goff[k]; e < S.gend[k]; ++e) {
        const int code = S.gent[e];
        const int64_t p = code > 0 ? code - 1 : -code - 1;
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] += epv[7 * p + c];
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) evat[7 * (size_t)k + c] = 0.5 * acc[c];
}

// ---------------------------------------------------------------- gather
__global__ __launch_bounds__(256) void k_gather(DevSystem S) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S.nall) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int e1 = S.gend[k];
    int e = S.goff[k];
    // eight entries per step: their pair forces are loaded together, then
    // added in list order (the same sums as one entry at a time)
    for (; e + 8 <= e1; e += 8) {
        double v[8][3], sg[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int code = S.gent[e + u];
            const int64_t p = code > 0 ? code - 1 : -code - 1;
            sg[u] = code > 0 ? 1.0 : -1.0;
            v[u][0] = S.gpair[3 * p];
            v[u][1] = S.gpair[3 * p + 1];
            v[u][2] = S.gpair[3 * p + 2];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            fx = fma(sg[u], v[u][0], fx);
            fy = fma(sg[u], v[u][1], fy);
            fz = fma(sg[u], v[u][2], fz);
        }
    }
    for (; e < e1; ++e) {
        const int code = S.gent[e];
        const int64_t p = code > 0 ? code - 1 : -code - 1;
        const double sg = code > 0 ? 1.0 : -1.0;
        fx = fma(sg, S.gpair[3 * p], fx);
        fy = fma(sg, S.gpair[3 * p + 1], fy);
        fz = fma(sg, S.gpair[3 * p + 2], fz);
    }
    S.f[3 * k] = fx;
    S.f[3 * k + 1] = fy;
    S.f[3 * k + 2] = fz;
}

// ---------------------------------------------------------------- energy / virial
__global__ __launch_bounds__(256) void k_energy_slot(DevModel M, DevSystem S) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= S.nblocks * kBlock) return;
    double e = 0.0;
    if (S.atom_of_slot[slot] >= 0 && M.pair_model) {
        e = S.Ep[slot];
    } else if (S.atom_of_slot[slot] >= 0) {
        e = M.n_p > 0 ? S.Ep[slot] : 0.0;
        for (int n = 0; n < M.n_fn; ++n) {
            if (M.q12_adj_levels) e += S.Eq12[(size_t)slot * M.n_fn + n];
            if (M.has_q34)
                for (int w = 0; w < kWavesPerWg; ++w) e += S.Eq34[((size_t)slot * M.n_fn + n) * kWavesPerWg + w];
        }
    }
    S.e_slot[slot] = e;
}

// Per-structure energy and virial of a batch: one workgroup per structure,
// fixed-order tree over its (contiguous) atoms.
__global__ __launch_bounds__(256) void k_struct_reduce(DevSystem S, const int32_t* sbeg, const int32_t* send,
                                                       double* out) {
    __shared__ double sh[256];
    const int st = blockIdx.x;
    const int a0 = sbeg[st], a1 = send[st];
    for (int c = 0; c < 7; ++c) {
        double v = 0.0;
        for (int i = a0 + threadIdx.x; i < a1; i += 256)
            v += c == 0 ? S.e_slot[S.slot_of_atom[i]] : S.vown[6 * (size_t)i + (c - 1)];
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[7 * st + c] = sh[0];
        __syncthreads();
    }
}

// Fixed-order two-level reduction: k_reduce1 sums contiguous chunks per
// workgroup, k_reduce2 sums the partials in one workgroup.
constexpr int kRedThreads = 256;

__global__ __launch_bounds__(kRedThreads) void k_reduce1(const double* in, int n, int stride, int ncomp,
                                                         double* part) {
    __shared__ double sh[kRedThreads];
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0;
        for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
            s += in[(size_t)i * stride + c];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = kRedThreads / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[c * gridDim.x + blockIdx.x] = sh[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRedThreads) void k_reduce2(const double* part, int nparts, int ncomp, double* out) {
    __shared__ double sh[kRedThreads];
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nparts; i += kRedThreads) s += part[c * nparts + i];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = kRedThreads / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = sh[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = 0.0;
}

__global__ __launch_bounds__(256) void k_copy(double4* __restrict__ dst, const double4* __restrict__ src, int64_t n) {
    for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}

}  // namespace



size_t lds_limit() {
    static int cached_dev = -1;
    static size_t cached = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (dev != cached_dev) {
        int a = 0, b = 0;
        if (hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) a = 0;
        if (hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) b = 0;
        cached = static_cast<size_t>(std::max(a, b));
        if (cached == 0) cached = 160 * 1024;  // CDNA4: 160 KB of LDS per workgroup
        cached_dev = dev;
    }
    return cached;
}

hipError_t launched(size_t lds_bytes) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return lds_bytes > lds_limit() ? hipErrorInvalidConfiguration : hipSuccess;
}

hipError_t launch_q12(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st) {
    if (s.nblocks == 0) return hipSuccess;
    if (feat) {
        // staged feature writes when the widest species' tile is small
        int wmax = 0;
        for (int t = 0; t < m.n_type && t < 2; ++t) wmax = std::max(wmax, m.prange[t][1] - m.prange[t][0]);
        const size_t lds = sizeof(double) * kQ12Pitch * wmax;
        const int stage = (wmax > 0 && lds <= kQ12MaxStage) ? 1 : 0;
        hipLaunchKernelGGL(k_q12<true>, dim3(s.nblocks, m.n_fn), dim3(64), stage ? lds : 0, st, m, s, stage);
    } else {
        if (!m.q12_adj_levels) return hipSuccess;  // every level runs in the order-3/4 kernel
        hipLaunchKernelGGL(k_q12<false>, dim3(s.nblocks, m.n_fn), dim3(64), 0, st, m, s, 0);
    }
    return launched();
}

hipError_t launch_poly(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot == 0 || m.n_p == 0) return hipSuccess;
    if (m.poly_fused) return launch_poly_sum(m, s, st);  // H and the partials came from the GEMM
    hipLaunchKernelGGL(k_poly, dim3((nslot + 3) / 4), dim3(256), 0, st, m, s);
    return launched();
}

hipError_t launch_poly_split(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot == 0 || m.n_p == 0) return hipSuccess;
    hipLaunchKernelGGL(k_poly_split, dim3((nslot + 3) / 4), dim3(256), 0, st, m, s);
    return launched();
}

hipError_t launch_atom_ev(const DevSystem& s, const double* epv, double* evat, hipStream_t st) {
    if (s.nall == 0) return hipSuccess;
    hipLaunchKernelGGL(k_atom_ev, dim3((s.nall + 255) / 256), dim3(256), 0, st, s, epv, evat);
    return launched();
}

hipError_t launch_gather(const DevModel&, const DevSystem& s, hipStream_t st) {
    if (s.nall == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((s.nall + 255) / 256), dim3(256), 0, st, s);
    return launched();
}

hipError_t launch_fill(double* dst, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(4096, (static_cast<int64_t>(n) + 255) / 256);
    hipLaunchKernelGGL(k_fill, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, dst, static_cast<int64_t>(n));
    return launched();
}

hipError_t launch_copy(void* dst, const void* src, int64_t nbytes, int reps, hipStream_t st) {
    const int64_t n = nbytes / static_cast<int64_t>(sizeof(double4));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, st, static_cast<double4*>(dst),
                           static_cast<const double4*>(src), n);
    return launched();
}

hipError_t launch_struct_reduce(const DevSystem& s, const int32_t* sbeg, const int32_t* send, int nstruct,
                                double* out, hipStream_t st) {
    if (nstruct <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_struct_reduce, dim3(nstruct), dim3(256), 0, st, s, sbeg, send, out);
    return launched();
}

hipError_t launch_energy_virial(const DevModel& m, const DevSystem& s, bool totals, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot > 0) hipLaunchKernelGGL(k_energy_slot, dim3((nslot + 255) / 256), dim3(256), 0, st, m, s);
    hipError_t e = launched();
    if (!totals) return e;
    const int parts = 120;
    hipLaunchKernelGGL(k_reduce1, dim3(parts), dim3(kRedThreads), 0, st, s.e_slot, nslot, 1, 1, s.red);
    e = first_error(e, launched());
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kRedThreads), 0, st, s.red, parts, 1, s.out);
    e = first_error(e, launched());
    hipLaunchKernelGGL(k_reduce1, dim3(parts), dim3(kRedThreads), 0, st, s.vown, s.nown, 6, 6, s.red + parts);
    e = first_error(e, launched());
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kRedThreads), 0, st, s.red + parts, parts, 6, s.out + 1);
    return first_error(e, launched());
}

}  // namespace mlip
