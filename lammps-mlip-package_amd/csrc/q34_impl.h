// Order-3/4 invariants of the gtinv model on CDNA4: polynomial features
// (FEAT = true) or the adjoint G = dE/da (FEAT = false).
//
// Workgroup = 8 waves on one (block of 64 same-type centres, radial index
// n).  The a_nlm tile (l <= lmax3, both species) sits in LDS as
// [s][lm][re,im][lane]; each wave owns whole (S, L) output groups and keeps
// G_{S,L,m<=0} in registers.  Reference: compute_anlm_uniq_products and the
// linear/polynomial adjoint terms of src/USER-MLIP/pair_mlip_gtinv.cpp
// (:260-343, :451-521), docs/formulation.md (adjoint G).
#pragma once

#include <hip/hip_runtime.h>

#include "device_api.h"
#include "device_common.h"

namespace mlip {
namespace q34 {

using namespace dev;

constexpr int kQ34Threads = 64 * kWavesPerWg;

// Workgroup = 8 waves on one (block of 64 same-type centres, n).  The a_nlm
// tile (l <= lmax3, both species) sits in LDS as [s][lm][re,im][lane]; each
// wave owns whole (S, L) output groups and keeps G_{S,L,m<=0} in registers.
struct Tile {
    const double* As;
    int nlmh3;
    int lane;
    __device__ __forceinline__ void get(int s, int l, int m, double& re, double& im) const {
        const int am = m < 0 ? -m : m;
        const int idx = s * nlmh3 + l * (l + 1) / 2 + l - am;
        re = As[(2 * idx) * kBlock + lane];
        im = As[(2 * idx + 1) * kBlock + lane];
        if (m > 0) {
            const double sg = (m & 1) ? -1.0 : 1.0;
            re *= sg;
            im *= -sg;
        }
    }
};

template <int L>
__device__ __forceinline__ void load_out_row(const Tile& T, int S, double (&lr)[L + 1], double (&li)[L + 1]) {
#pragma unroll
    for (int k = 0; k <= L; ++k) T.get(S, L, -k, lr[k], li[k]);
}

template <int L, bool DOG, bool DOD>
__device__ __forceinline__ void visit3(const Tile& T, const DevModel& M, const VisitQ34& v, int S, double H,
                                       double (&gr)[L + 1], double (&gi)[L + 1], double& d) {
    double lr[L + 1], li[L + 1];
    if (DOD) load_out_row<L>(T, S, lr, li);
    const int la = uniform(v.la), lb = uniform(v.lb), sa = uniform(v.sa), sb = uniform(v.sb);
    cptr<double> cg = as_const(M.coef) + uniform(v.coef);
    cptr<double> cd = as_const(M.coef) + uniform(v.dcoef);
    for (int ia = 0; ia <= 2 * la; ++ia) {
        const int ma = ia - la;
        double xr, xi;
        T.get(sa, la, ma, xr, xi);
        cptr<double> cgk = cg + ia * (L + 1);
        cptr<double> cdk = cd + ia * (L + 1);
#pragma unroll
        for (int k = 0; k <= L; ++k) {
            const int mb = k - ma;
            if (mb < -lb || mb > lb) continue;
            double yr, yi;
            T.get(sb, lb, mb, yr, yi);
            const double pr = xr * yr - xi * yi;
            const double pi = xr * yi + xi * yr;
            if (DOG) {
                const double hc = H * cgk[k];
                gr[k] = fma(hc, pr, gr[k]);
                gi[k] = fma(hc, pi, gi[k]);
            }
            if (DOD) d = fma(cdk[k], lr[k] * pr - li[k] * pi, d);
        }
    }
}

template <int L, bool DOG, bool DOD>
__device__ __forceinline__ void visit4(const Tile& T, const DevModel& M, const VisitQ34& v, int S, double H,
                                       double (&gr)[L + 1], double (&gi)[L + 1], double& d) {
    double lr[L + 1], li[L + 1];
    if (DOD) load_out_row<L>(T, S, lr, li);
    const int la = uniform(v.la), lb = uniform(v.lb), lc = uniform(v.lc);
    const int sa = uniform(v.sa), sb = uniform(v.sb), sc = uniform(v.sc);
    cptr<double> cg = as_const(M.coef) + uniform(v.coef);
    cptr<double> cd = as_const(M.coef) + uniform(v.dcoef);
    const int nb = 2 * lb + 1;
    for (int ia = 0; ia <= 2 * la; ++ia) {
        const int ma = ia - la;
        double xr, xi;
        T.get(sa, la, ma, xr, xi);
        for (int ib = 0; ib < nb; ++ib) {
            const int mb = ib - lb;
            double yr, yi;
            T.get(sb, lb, mb, yr, yi);
            const double qr = xr * yr - xi * yi;
            const double qi = xr * yi + xi * yr;
            cptr<double> cgk = cg + (ia * nb + ib) * (L + 1);
            cptr<double> cdk = cd + (ia * nb + ib) * (L + 1);
#pragma unroll
            for (int k = 0; k <= L; ++k) {
                const int mc = k - ma - mb;
                if (mc < -lc || mc > lc) continue;
                double zr, zi;
                T.get(sc, lc, mc, zr, zi);
                const double pr = qr * zr - qi * zi;
                const double pi = qr * zi + qi * zr;
                if (DOG) {
                    const double hc = H * cgk[k];
                    gr[k] = fma(hc, pr, gr[k]);
                    gi[k] = fma(hc, pi, gi[k]);
                }
                if (DOD) d = fma(cdk[k], lr[k] * pr - li[k] * pi, d);
            }
        }
    }
}

// Fully unrolled visits for the signatures of q34_signatures.inc.  The b
// (and c) rows are loaded once per visit into registers, with the m > 0
// entries formed by the static conjugation a_{l,m} = (-1)^m conj(a_{l,-m});
// the a factor is read once per m_a.  Coefficients are compact (valid terms
// in loop order), so every coefficient offset is a compile-time constant.
template <int LB>
__device__ __forceinline__ void load_row(const Tile& T, int s, double (&r)[2 * LB + 1], double (&i)[2 * LB + 1]) {
#pragma unroll
    for (int k = 0; k <= LB; ++k) {
        double re, im;
        T.get(s, LB, -k, re, im);
        r[LB - k] = re;
        i[LB - k] = im;
        if (k > 0) {
            const double sg = (k & 1) ? -1.0 : 1.0;
            r[LB + k] = sg * re;
            i[LB + k] = -sg * im;
        }
    }
}

template <int L, int LA, int LB, bool DOG, bool DOD>
__device__ __forceinline__ void visit3u(const Tile& T, int S, int sa, int sb, cptr<double> cg,
                                        cptr<double> cd, double H, double (&gr)[L + 1],
                                        double (&gi)[L + 1], double& d) {
    double br[2 * LB + 1], bi[2 * LB + 1];
    load_row<LB>(T, sb, br, bi);
    int t = 0;
#pragma unroll
    for (int ia = 0; ia <= 2 * LA; ++ia) {
        const int ma = ia - LA;
        double xr, xi;
        T.get(sa, LA, ma, xr, xi);
        // d-forming visits read the output factor a_{S,L,-k} term by term; the
        // barrier keeps the compiler from hoisting those reads into registers
        if (DOD) asm volatile("" ::: "memory");
#pragma unroll
        for (int k = 0; k <= L; ++k) {
            const int mb = k - ma;
            if (mb < -LB || mb > LB) continue;
            const double yr = br[mb + LB], yi = bi[mb + LB];
            const double pr = xr * yr - xi * yi;
            const double pi = xr * yi + xi * yr;
            if (DOG) {
                const double hc = H * cg[t];
                gr[k] = fma(hc, pr, gr[k]);
                gi[k] = fma(hc, pi, gi[k]);
            }
            if (DOD) {
                double lr, li;
                T.get(S, L, -k, lr, li);
                d = fma(cd[t], lr * pr - li * pi, d);
            }
            ++t;
        }
    }
}

template <int L, int LA, int LB, int LC, bool DOG, bool DOD>
__device__ __forceinline__ void visit4u(const Tile& T, int S, int sa, int sb, int sc, cptr<double> cg,
                                        cptr<double> cd, double H, double (&gr)[L + 1],
                                        double (&gi)[L + 1], double& d) {
    double br[2 * LB + 1], bi[2 * LB + 1], cr[2 * LC + 1], ci[2 * LC + 1];
    load_row<LB>(T, sb, br, bi);
    load_row<LC>(T, sc, cr, ci);
    int t = 0;
#pragma unroll
    for (int ia = 0; ia <= 2 * LA; ++ia) {
        const int ma = ia - LA;
        double xr, xi;
        T.get(sa, LA, ma, xr, xi);
#pragma unroll
        for (int ib = 0; ib <= 2 * LB; ++ib) {
            const int mb = ib - LB;
            const double qr = xr * br[ib] - xi * bi[ib];
            const double qi = xr * bi[ib] + xi * br[ib];
            if (DOD) asm volatile("" ::: "memory");
#pragma unroll
            for (int k = 0; k <= L; ++k) {
                const int mc = k - ma - mb;
                if (mc < -LC || mc > LC) continue;
                const double zr = cr[mc + LC], zi = ci[mc + LC];
                const double pr = qr * zr - qi * zi;
                const double pi = qr * zi + qi * zr;
                if (DOG) {
                    const double hc = H * cg[t];
                    gr[k] = fma(hc, pr, gr[k]);
                    gi[k] = fma(hc, pi, gi[k]);
                }
                if (DOD) {
                    double lr, li;
                    T.get(S, L, -k, lr, li);
                    d = fma(cd[t], lr * pr - li * pi, d);
                }
                ++t;
            }
        }
    }
}

#include "q34_signatures.inc"

// Dispatch on (l_a, l_b[, l_c]) for a fixed output L; returns false for a
// signature outside the unrolled set (the caller then runs the general loop).
template <int L, bool DOG, bool DOD>
struct Unrolled;

#define MLIP_CASE3(A, B) \
    case (A) * 16 + (B): visit3u<LV, A, B, DOG, DOD>(T, S, v.sa, v.sb, cg, cd, H, gr, gi, d); return true;
#define MLIP_CASE4(A, B, C)                                                                       \
    case ((A) * 16 + (B)) * 16 + (C):                                                             \
        visit4u<LV, A, B, C, DOG, DOD>(T, S, v.sa, v.sb, v.sc, cg, cd, H, gr, gi, d); return true;
#define MLIP_UNROLLED(LL, CASES3, CASES4)                                                                     \
    template <bool DOG, bool DOD>                                                                             \
    struct Unrolled<LL, DOG, DOD> {                                                                           \
        static constexpr int LV = LL;                                                                         \
        __device__ static bool run(const Tile& T, const VisitQ34& v, int S, int order, int la, int lb, int lc, \
                                   cptr<double> cg, cptr<double> cd, double H, double (&gr)[LL + 1],          \
                                   double (&gi)[LL + 1], double& d) {                                         \
            if (order == 3) {                                                                                 \
                switch (la * 16 + lb) { CASES3 default: return false; }                                       \
            }                                                                                                 \
            switch ((la * 16 + lb) * 16 + lc) { CASES4 default: return false; }                              \
        }                                                                                                     \
    };
#define MLIP_NO_CASES
MLIP_UNROLLED(0, MLIP_SIG3_L0(MLIP_CASE3), MLIP_SIG4_L0(MLIP_CASE4))
MLIP_UNROLLED(1, MLIP_SIG3_L1(MLIP_CASE3), MLIP_SIG4_L1(MLIP_CASE4))
MLIP_UNROLLED(2, MLIP_SIG3_L2(MLIP_CASE3), MLIP_SIG4_L2(MLIP_CASE4))
MLIP_UNROLLED(3, MLIP_SIG3_L3(MLIP_CASE3), MLIP_NO_CASES)
MLIP_UNROLLED(4, MLIP_SIG3_L4(MLIP_CASE3), MLIP_NO_CASES)
MLIP_UNROLLED(5, MLIP_SIG3_L5(MLIP_CASE3), MLIP_NO_CASES)
MLIP_UNROLLED(6, MLIP_SIG3_L6(MLIP_CASE3), MLIP_NO_CASES)
MLIP_UNROLLED(7, MLIP_SIG3_L7(MLIP_CASE3), MLIP_NO_CASES)
MLIP_UNROLLED(8, MLIP_SIG3_L8(MLIP_CASE3), MLIP_NO_CASES)
template <int L, bool DOG, bool DOD>
struct Unrolled {
    __device__ static bool run(const Tile&, const VisitQ34&, int, int, int, int, int, cptr<double>, cptr<double>,
                               double, double (&)[L + 1], double (&)[L + 1], double&) {
        return false;
    }
};

template <int L, bool FEAT>
__device__ __forceinline__ void q34_group(const Tile& T, const DevModel& M, const DevSystem& S, const GroupQ34& g, int blk,
                          int n, double& eacc) {
    double gr[L + 1], gi[L + 1];
    const int Sx = uniform(g.S);
#pragma unroll
    for (int k = 0; k <= L; ++k) gr[k] = gi[k] = 0.0;
    const int slot = blk * kBlock + T.lane;
    cptr<VisitQ34> vis = as_const(FEAT ? M.vfeat : M.vadj);
    const int vb = uniform(g.vbeg), ve = uniform(g.vend);
    for (int vi = vb; vi < ve; ++vi) {
        const VisitQ34 v = vis[vi];
        const int order = uniform(v.order), flags = uniform(v.flags), pslot = uniform(v.pslot);
        const int itc = uniform(v.itc);
        double d = 0.0;
        const bool fast = (flags & 4) != 0;
        const int la = uniform(v.la), lb = uniform(v.lb), lc = uniform(v.lc);
        cptr<double> cg = as_const(M.coef) + uniform(v.coef);
        cptr<double> cd = as_const(M.coef) + uniform(v.dcoef);
        if (FEAT) {
            if (!(fast && Unrolled<L, false, true>::run(T, v, Sx, order, la, lb, lc, cg, cd, 0.0, gr, gi, d))) {
                if (order == 3)
                    visit3<L, false, true>(T, M, v, Sx, 0.0, gr, gi, d);
                else
                    visit4<L, false, true>(T, M, v, Sx, 0.0, gr, gi, d);
            }
            S.D[(size_t)slot * M.n_p + n * M.n_pitc + pslot] = d;
            continue;
        }
        const double H =
            pslot >= 0 ? S.H[(size_t)slot * M.n_p + n * M.n_pitc + pslot] : as_const(M.weight)[n * M.n_gtinv + itc];
        if (flags & 1) {
            if (!(fast && Unrolled<L, true, true>::run(T, v, Sx, order, la, lb, lc, cg, cd, H, gr, gi, d))) {
                if (order == 3)
                    visit3<L, true, true>(T, M, v, Sx, H, gr, gi, d);
                else
                    visit4<L, true, true>(T, M, v, Sx, H, gr, gi, d);
            }
            eacc = fma(as_const(M.weight)[n * M.n_gtinv + itc], d, eacc);
        } else {
            if (!(fast && Unrolled<L, true, false>::run(T, v, Sx, order, la, lb, lc, cg, cd, H, gr, gi, d))) {
                if (order == 3)
                    visit3<L, true, false>(T, M, v, Sx, H, gr, gi, d);
                else
                    visit4<L, true, false>(T, M, v, Sx, H, gr, gi, d);
            }
        }
    }
    if (!FEAT) {
#pragma unroll
        for (int k = 0; k <= L; ++k) {
            const size_t ir = a_index(M, blk, n, Sx, lmh(L, -k), 0, T.lane);
            const size_t ii = a_index(M, blk, n, Sx, lmh(L, -k), 1, T.lane);
            S.G[ir] += gr[k];
            S.G[ii] += gi[k];
        }
    }
}


template <bool FEAT>
__global__ __launch_bounds__(kQ34Threads) void k_q34(DevModel M, DevSystem S) {
    extern __shared__ double As[];
    const int blk = blockIdx.x, n = blockIdx.y;
    const int t = S.block_type[blk];
    const int per_s = M.n_lmh3 * 2 * kBlock;
    for (int s = 0; s < M.n_type; ++s) {
        const double* src = S.A + a_index(M, blk, n, s, 0, 0, 0);
        for (int k = threadIdx.x; k < per_s; k += kQ34Threads) As[s * per_s + k] = src[k];
    }
    __syncthreads();
    const int wave = uniform(threadIdx.x >> 6);
    Tile T{As, M.n_lmh3, static_cast<int>(threadIdx.x & 63)};
    cptr<GroupQ34> groups = as_const(FEAT ? M.gfeat : M.gadj);
    const int gb = FEAT ? M.gfeat_beg[t] : M.gadj_beg[t];
    const int ge = FEAT ? M.gfeat_end[t] : M.gadj_end[t];
    double eacc = 0.0;
    for (int gi = gb; gi < ge; ++gi) {
        const GroupQ34 g = groups[gi];
        if (uniform(g.wave) != wave) continue;
        switch (uniform(g.L)) {
            case 0: q34_group<0, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 1: q34_group<1, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 2: q34_group<2, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 3: q34_group<3, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 4: q34_group<4, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 5: q34_group<5, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 6: q34_group<6, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 7: q34_group<7, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 8: q34_group<8, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 9: q34_group<9, FEAT>(T, M, S, g, blk, n, eacc); break;
            case 10: q34_group<10, FEAT>(T, M, S, g, blk, n, eacc); break;
            default: break;
        }
    }
    if (!FEAT) S.Eq34[((size_t)(blk * kBlock + T.lane) * M.n_fn + n) * kWavesPerWg + wave] = eacc;
}


}  // namespace q34
}  // namespace mlip
