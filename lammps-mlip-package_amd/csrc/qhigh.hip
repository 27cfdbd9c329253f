// Order-5/6 invariants of the gtinv model (the table's order-5 and order-6
// families; every factor has l <= 1 there, the kernel takes up to l = 4):
// polynomial features (FEAT) or their share of the adjoint G = dE/da and of
// the energy of linear families.  Reference: the same terms of
// compute_anlm_uniq_products / compute_anlm_linear_term /
// compute_anlm_polynomial_model_correction (src/USER-MLIP/
// pair_mlip_gtinv.cpp:260-343, 451-521) for irreps of order 5 and 6.
//
// Workgroup = one wave on (block of 64 same-species centres, radial index n),
// lane = centre.  The centre's a_{s,l<=lmaxh,m<=0} is staged in LDS (each
// lane reads only its own entries); a visit is one (family, factor position)
// with its terms listed: G_{S,L,-k} += H sum_terms C prod(other factors),
// and for linear families the energy share (1/q) sum_k w_k Re(G_k a_{S,L,-k})
// (Euler's theorem summed over the q factor positions), added to the
// order-3/4 energy slot 0 of the centre after the order-3/4 kernel.
#include <hip/hip_runtime.h>

#include "device_api.h"
#include "device_common.h"

namespace mlip {

namespace {

using namespace dev;

struct HiTile {
    const double2* a;  // [s][lm][lane]
    int nh, lane;
    __device__ __forceinline__ void get(int s, int l, int m, double& re, double& im) const {
        const int am = m < 0 ? -m : m;
        const double2 v = a[(s * nh + l * (l + 1) / 2 + l - am) * kBlock + lane];
        re = v.x;
        im = v.y;
        if (m > 0) {  // a_{l,m} = (-1)^m conj(a_{l,-m})
            const double sg = (m & 1) ? -1.0 : 1.0;
            re *= sg;
            im *= -sg;
        }
    }
};

template <bool FEAT>
__global__ __launch_bounds__(64) void k_qhigh(DevModel M, DevSystem S) {
    extern __shared__ double2 hi_tile[];
    const int blk = blockIdx.x, n = blockIdx.y, lane = threadIdx.x;
    const int t = S.block_type[blk];
    const int nh = M.n_lmhh;
    for (int s = 0; s < M.n_type; ++s)
        for (int lm = 0; lm < nh; ++lm)
            hi_tile[(s * nh + lm) * kBlock + lane] =
                double2{S.A[a_index(M, blk, n, s, lm, 0, lane)], S.A[a_index(M, blk, n, s, lm, 1, lane)]};
    const HiTile T{hi_tile, nh, lane};
    const int slot = blk * kBlock + lane;
    const int64_t prow = (int64_t)slot * M.n_pr + pcol(M, t, n, 0);
    const int vb = FEAT ? M.hfeat_beg[t] : M.hadj_beg[t];
    const int ve = FEAT ? M.hfeat_end[t] : M.hadj_end[t];
    cptr<VisitHi> vis = as_const(M.hvis);
    cptr<int32_t> term = as_const(M.hterm);
    cptr<double> coef = as_const(M.hcoef);
    double eacc = 0.0;
    for (int vi = vb; vi < ve; ++vi) {
        const VisitHi v = vis[vi];
        const int L = uniform(v.L), Sx = uniform(v.S), q = uniform(v.order), pslot = uniform(v.pslot);
        const int no = q - 1;
        double gr[kMaxLHi + 1], gi[kMaxLHi + 1];
#pragma unroll
        for (int k = 0; k <= kMaxLHi; ++k) gr[k] = gi[k] = 0.0;
        double d = 0.0;
        for (int ti = uniform(v.tbeg); ti < uniform(v.tend); ++ti) {
            const int code = term[ti];
            const double c = coef[ti];
            const int k = code & 15;
            double pr = 1.0, pi = 0.0;
            for (int o = 0; o < no; ++o) {
                const int ls = uniform(v.ls[o]);
                const int mo = ((code >> (4 + 4 * o)) & 15) - 8;
                double yr, yi;
                T.get(ls >> 8, ls & 255, mo, yr, yi);
                const double tr = pr * yr - pi * yi;
                pi = pr * yi + pi * yr;
                pr = tr;
            }
            if (FEAT) {
                double ar, ai;
                T.get(Sx, L, -k, ar, ai);
                d = fma(c, ar * pr - ai * pi, d);
            } else {
#pragma unroll
                for (int kk = 0; kk <= kMaxLHi; ++kk)
                    if (kk == k) {
                        gr[kk] = fma(c, pr, gr[kk]);
                        gi[kk] = fma(c, pi, gi[kk]);
                    }
            }
        }
        if (FEAT) {
            S.D[prow + pslot] = d;
            continue;
        }
        const double H = pslot >= 0 ? S.H[prow + pslot] : as_const(M.weight)[n * M.n_gtinv + uniform(v.itc)];
        double dot = 0.0;
#pragma unroll
        for (int k = 0; k <= kMaxLHi; ++k) {
            if (k > L) break;
            const double hr = H * gr[k], hi = H * gi[k];
            double* g = S.G + a_index(M, blk, n, Sx, lmh(L, -k), 0, lane);
            g[0] += hr;
            g[kBlock] += hi;  // imaginary part: next component, same lane
            double ar, ai;
            T.get(Sx, L, -k, ar, ai);
            dot = fma(k == 0 ? 1.0 : 2.0, hr * ar - hi * ai, dot);
        }
        if (pslot < 0) eacc = fma(1.0 / q, dot, eacc);
    }
    if (!FEAT) S.Eq34[((size_t)slot * M.n_fn + n) * kWavesPerWg] += eacc;
}

}  // namespace

hipError_t launch_qhigh(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st) {
    if (s.nblocks == 0 || m.lmaxh < 0) return hipSuccess;
    const size_t lds = sizeof(double2) * m.n_type * m.n_lmhh * kBlock;
    if (lds > lds_limit()) return hipErrorInvalidConfiguration;
    if (feat)
        hipLaunchKernelGGL(k_qhigh<true>, dim3(s.nblocks, m.n_fn), dim3(64), lds, st, m, s);
    else
        hipLaunchKernelGGL(k_qhigh<false>, dim3(s.nblocks, m.n_fn), dim3(64), lds, st, m, s);
    return launched(lds);
}

}  // namespace mlip
