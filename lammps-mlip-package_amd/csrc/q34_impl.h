// Order-3/4 invariants of the gtinv model on CDNA4: polynomial features
// (FEAT = true) or the adjoint G = dE/da (FEAT = false).
//
// Workgroup = kWavesPerWg waves on one (block of 64 same-type centres,
// radial index n).  The a_nlm tile (l <= lmax3, both species) sits in LDS as
// [s][lm][lane](re, im); each wave owns (S, L) output groups (or balanced
// parts of heavy ones) and keeps G_{S,L,m<=0} in registers.  For L <=
// kPairMaxL the rows of both species form one group, whose visits common
// to the two rows form their products once (q34_group_pair).  Visits with
// fully unrolled signatures run in one kernel, the rest in a second
// general-loop kernel (only launched for models that have such visits), so
// the common kernel's register budget allows more waves per SIMD.  Reference: compute_anlm_uniq_products and the
// linear/polynomial adjoint terms of src/USER-MLIP/pair_mlip_gtinv.cpp
// (:260-343, :451-521), docs/formulation.md (adjoint G).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>

#include "device_api.h"
#include "device_common.h"

namespace mlip {
namespace q34 {

using namespace dev;

constexpr int kQ34Threads = 64 * kWavesPerWg;

struct Tile {
    const double2* As;  // [s][lm][lane] (re, im): one 16-byte LDS read per entry
    int nlmh3;
    int lane;
    const double* Ag;   // a_nlm of the (block, n) in HBM: [s][lm][re, im][lane]
    int nlmh;
    // the row (s, l) of a level above the LDS tile (fused order-1/2 rows
    // l > lmax3) in HBM: entry m <= 0 at (l - |m|) 2 kBlock from the
    // returned base.  The base is opaque, so a visit keeps one pointer rather
    // than every entry's address hoisted out of the visit loop.
    __device__ __forceinline__ const double* rowg(int s, int l) const {
        const double* p = Ag + (static_cast<size_t>(s * nlmh + l * (l + 1) / 2) * 2) * kBlock + lane;
        asm volatile("" : "+v"(p));
        return p;
    }
    __device__ __forceinline__ static void entryg(const double* row, int l, int m, double& re, double& im) {
        const int am = m < 0 ? -m : m;
        re = row[(2 * (l - am)) * kBlock];
        im = row[(2 * (l - am) + 1) * kBlock];
        if (m > 0) {
            const double sg = (m & 1) ? -1.0 : 1.0;
            re *= sg;
            im *= -sg;
        }
    }
    __device__ __forceinline__ void get(int s, int l, int m, double& re, double& im) const {
        const int am = m < 0 ? -m : m;
        const int idx = s * nlmh3 + l * (l + 1) / 2 + l - am;
        const double2 v = As[idx * kBlock + lane];
        re = v.x;
        im = v.y;
        if (m > 0) {
            const double sg = (m & 1) ? -1.0 : 1.0;
            re *= sg;
            im *= -sg;
        }
    }
};

// A visit decoded from its packed device form (wave-uniform fields).
struct VDec {
    int order, itc, la, sa, lb, sb, lc, sc, pslot, flags, coef, dcoef;
};

__device__ __forceinline__ VDec decode(const VisitPack& p) {
    const int s = uniform(p.sig);
    VDec v;
    v.order = s & 15;
    v.la = (s >> 4) & 15;
    v.lb = (s >> 8) & 15;
    v.lc = (s >> 12) & 15;
    v.sa = (s >> 16) & 15;
    v.sb = (s >> 20) & 15;
    v.sc = (s >> 24) & 15;
    v.flags = (s >> 28) & 15;
    v.coef = uniform(p.coef);
    v.dcoef = v.coef;
    v.pslot = uniform(p.pslot);
    v.itc = uniform(p.itc);
    return v;
}

template <int L>
__device__ __forceinline__ void load_out_row(const Tile& T, int S, double (&lr)[L + 1], double (&li)[L + 1]) {
#pragma unroll
    for (int k = 0; k <= L; ++k) T.get(S, L, -k, lr[k], li[k]);
}

template <int L, bool DOG, bool DOD>
__device__ __forceinline__ void visit3(const Tile& T, const DevModel& M, const VDec& v, int S, double H,
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
        if (DOG) {
            xr *= H;
            xi *= H;
        }
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
                gr[k] = fma(cgk[k], pr, gr[k]);
                gi[k] = fma(cgk[k], pi, gi[k]);
            }
            if (DOD) d = fma(cdk[k], lr[k] * pr - li[k] * pi, d);
        }
    }
}

template <int L, bool DOG, bool DOD>
__device__ __forceinline__ void visit4(const Tile& T, const DevModel& M, const VDec& v, int S, double H,
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
        if (DOG) {
            xr *= H;
            xi *= H;
        }
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
                    gr[k] = fma(cgk[k], pr, gr[k]);
                    gi[k] = fma(cgk[k], pi, gi[k]);
                }
                if (DOD) d = fma(cdk[k], lr[k] * pr - li[k] * pi, d);
            }
        }
    }
}

// Fully unrolled visits for the signatures of q34_signatures.inc.  The b
// (and c) rows are loaded once per visit into registers, m <= 0 entries
// only: an m > 0 factor is a_{l,m} = (-1)^m conj(a_{l,-m}), whose sign and
// conjugation fold into the (fully unrolled, so compile-time) product.  The
// a factor is read once per m_a.  Coefficients are compact (valid terms in
// loop order), so every coefficient offset is a compile-time constant.
template <int LB>
__device__ __forceinline__ void load_row(const Tile& T, int s, double (&r)[LB + 1], double (&i)[LB + 1]) {
#pragma unroll
    for (int k = 0; k <= LB; ++k) T.get(s, LB, -k, r[k], i[k]);
}

// a_{l,m} from a stored row of m <= 0 entries (m is a compile-time value
// after unrolling, so the selects fold away)
template <int LB>
__device__ __forceinline__ void factor(const double (&r)[LB + 1], const double (&i)[LB + 1], int m, double& yr,
                                       double& yi) {
    const int am = m < 0 ? -m : m;
    const double re = r[am], im = i[am];
    if (m <= 0) {
        yr = re;
        yi = im;
    } else if (m & 1) {
        yr = -re;
        yi = im;
    } else {
        yr = re;
        yi = -im;
    }
}

// SC: the a factor carries H (G accumulation); without it the visit adds
// the unscaled products into a visit accumulator (species-paired groups)
template <int L, int LA, int LB, bool DOG, bool DOD, bool SC>
__device__ __forceinline__ void visit3u(const Tile& T, int S, int sa, int sb, cptr<double> cg,
                                        cptr<double> cd, double H, double (&gr)[L + 1],
                                        double (&gi)[L + 1], double& d) {
    double br[LB + 1], bi[LB + 1];
    load_row<LB>(T, sb, br, bi);
    int t = 0;
    double xr, xi;
    T.get(sa, LA, -LA, xr, xi);
#pragma unroll
    for (int ia = 0; ia <= 2 * LA; ++ia) {
        const int ma = ia - LA;
        // next a factor requested one m_a ahead
        double nxr = 0.0, nxi = 0.0;
        if (ia < 2 * LA) T.get(sa, LA, ma + 1, nxr, nxi);
        if (DOG && SC) {
            xr *= H;
            xi *= H;
        }
        // d-forming visits read the output factor a_{S,L,-k} term by term; the
        // barrier keeps the compiler from hoisting those reads into registers
        if (DOD) asm volatile("" ::: "memory");
#pragma unroll
        for (int k = 0; k <= L; ++k) {
            const int mb = k - ma;
            if (mb < -LB || mb > LB) continue;
            double yr, yi;
            factor<LB>(br, bi, mb, yr, yi);
            const double pr = xr * yr - xi * yi;
            const double pi = xr * yi + xi * yr;
            if (DOG) {
                gr[k] = fma(cg[t], pr, gr[k]);
                gi[k] = fma(cg[t], pi, gi[k]);
            }
            if (DOD) {
                double lr, li;
                T.get(S, L, -k, lr, li);
                d = fma(cd[t], lr * pr - li * pi, d);
            }
            ++t;
        }
        // one m_a per scheduling region: interleaving all of them costs more
        // registers (and occupancy) than it hides latency; scalar
        // instructions (the coefficient loads) may still cross, so the next
        // region's loads issue early and merge into wider ones
        __builtin_amdgcn_sched_barrier(4);
        xr = nxr;
        xi = nxi;
    }
}

template <int L, int LA, int LB, int LC, bool DOG, bool DOD, bool SC>
__device__ __forceinline__ void visit4u(const Tile& T, int S, int sa, int sb, int sc, cptr<double> cg,
                                        cptr<double> cd, double H, double (&gr)[L + 1],
                                        double (&gi)[L + 1], double& d) {
    double br[LB + 1], bi[LB + 1], cr[LC + 1], ci[LC + 1];
    load_row<LB>(T, sb, br, bi);
    load_row<LC>(T, sc, cr, ci);
    int t = 0;
#pragma unroll
    for (int ia = 0; ia <= 2 * LA; ++ia) {
        const int ma = ia - LA;
        double xr, xi;
        T.get(sa, LA, ma, xr, xi);
        if (DOG && SC) {
            xr *= H;
            xi *= H;
        }
#pragma unroll
        for (int ib = 0; ib <= 2 * LB; ++ib) {
            const int mb = ib - LB;
            double yr, yi;
            factor<LB>(br, bi, mb, yr, yi);
            const double qr = xr * yr - xi * yi;
            const double qi = xr * yi + xi * yr;
            if (DOD) asm volatile("" ::: "memory");
#pragma unroll
            for (int k = 0; k <= L; ++k) {
                const int mc = k - ma - mb;
                if (mc < -LC || mc > LC) continue;
                double zr, zi;
                factor<LC>(cr, ci, mc, zr, zi);
                const double pr = qr * zr - qi * zi;
                const double pi = qr * zi + qi * zr;
                if (DOG) {
                    gr[k] = fma(cg[t], pr, gr[k]);
                    gi[k] = fma(cg[t], pi, gi[k]);
                }
                if (DOD) {
                    double lr, li;
                    T.get(S, L, -k, lr, li);
                    d = fma(cd[t], lr * pr - li * pi, d);
                }
                ++t;
            }
            __builtin_amdgcn_sched_barrier(4);
        }
    }
}

#include "q34_signatures.inc"

// Dispatch on (l_a, l_b[, l_c]) for a fixed output L; returns false for a
// signature outside the unrolled set (the caller then runs the general loop).
template <int L, bool DOG, bool DOD, bool SC>
struct Unrolled;

#define MLIP_CASE3(A, B) \
    case (A) * 16 + (B): visit3u<LV, A, B, DOG, DOD, SC>(T, S, v.sa, v.sb, cg, cd, H, gr, gi, d); return true;
#define MLIP_CASE4(A, B, C)                                                                       \
    case ((A) * 16 + (B)) * 16 + (C):                                                             \
        visit4u<LV, A, B, C, DOG, DOD, SC>(T, S, v.sa, v.sb, v.sc, cg, cd, H, gr, gi, d); return true;
#define MLIP_UNROLLED(LL, CASES3, CASES4)                                                                     \
    template <bool DOG, bool DOD, bool SC>                                                                    \
    struct Unrolled<LL, DOG, DOD, SC> {                                                                       \
        static constexpr int LV = LL;                                                                         \
        __device__ static bool run(const Tile& T, const VDec& v, int S, int order, int la, int lb, int lc,     \
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
template <int L, bool DOG, bool DOD, bool SC>
struct Unrolled {
    __device__ static bool run(const Tile&, const VDec&, int, int, int, int, int, cptr<double>, cptr<double>,
                               double, double (&)[L + 1], double (&)[L + 1], double&) {
        return false;
    }
};

// sum_{m <= 0} w_m Re(G_m a_{S,L,m}) over the accumulated row (w_0 = 1,
// w_{m<0} = 2): the derivative of the row's share of the energy along
// a -> (1 + eps) a.  For a homogeneous family of order q summed over all its
// factor positions this is q times its energy (Euler).
template <int L>
__device__ __forceinline__ double row_dot(const Tile& T, int S, const double (&gr)[L + 1],
                                          const double (&gi)[L + 1], bool glob = false) {
    double sum = 0.0;
    const double* rg = glob ? T.rowg(S, L) : nullptr;
#pragma unroll
    for (int k = 0; k <= L; ++k) {
        double ar, ai;
        if (glob)
            Tile::entryg(rg, L, -k, ar, ai);
        else
            T.get(S, L, -k, ar, ai);
        sum = fma(k == 0 ? 1.0 : 2.0, gr[k] * ar - gi[k] * ai, sum);
    }
    return sum;
}

// &G[blk, n, s = 0, lm = 0, re, lane]; entry (s, lm, c) lies at
// ((s n_lmh + lm) 2 + c) kBlock from it.  Formed after a group's visits and
// made opaque, so that the compiler keeps one pointer instead of hoisting
// every entry's address out of the visit loop (where they would hold
// registers across all visits).
__device__ __forceinline__ double* g_row_base(const DevModel& M, const DevSystem& S, int blk, int n, int lane) {
    double* p = S.G + a_index(M, blk, n, 0, 0, 0, lane);
    asm volatile("" : "+v"(p));
    return p;
}

// H of an adjoint visit from its packed source h: the polynomial adjoint at
// rank h of the slot's compact row (h >= 0), else the linear weight of
// family -1 - h at radial index n
__device__ __forceinline__ double h_value(const DevModel& M, const DevSystem& S, int64_t prow, int n, int h) {
    return h >= 0 ? S.H[prow + h] : as_const(M.weight)[n * M.n_gtinv + (-1 - h)];
}

// energy weight of a linear class (order q = c + 1): 1 / q of the row's dot
// product gained across it
__device__ __forceinline__ double class_weight(int c) {
    return c == 0 ? 1.0 : c == 1 ? 0.5 : c == 2 ? (1.0 / 3.0) : 0.25;
}

// fused order-1/2 adjoint term into one row: G_{L,-k} += H C_k a_{s_a,L,+k}
// (order 1: G_{0,0} += H C)
template <int L>
__device__ __forceinline__ void fused12(const Tile& T, int order, int sa, double H, cptr<double> c1,
                                        double (&gr)[L + 1], double (&gi)[L + 1]) {
    if (order == 1) {
        gr[0] = fma(H, c1[0], gr[0]);
        return;
    }
#pragma unroll
    for (int k = 0; k <= L; ++k) {
        double yr, yi;
        T.get(sa, L, k, yr, yi);
        const double hc = H * c1[k];
        gr[k] = fma(hc, yr, gr[k]);
        gi[k] = fma(hc, yi, gi[k]);
    }
}

// Species-paired adjoint group: rows G_{0,L} and G_{1,L} in registers.  A
// visit's products (unscaled) go into the visit accumulator t, which is then
// added into its row with its family's H, or into both rows with the two
// families' H where the visit is shared; fused order-1/2 visits add into
// their row directly.  Each row keeps its own energy classes.
template <int L>
__device__ __forceinline__ void q34_group_pair(const Tile& T, const DevModel& M, const DevSystem& S,
                                               const GroupQ34& g, int blk, int n, double* red, double& eacc) {
    double g0r[L + 1], g0i[L + 1], g1r[L + 1], g1i[L + 1];
#pragma unroll
    for (int k = 0; k <= L; ++k) g0r[k] = g0i[k] = g1r[k] = g1i[k] = 0.0;
    const int slot = blk * kBlock + T.lane;
    const int tsp = uniform(S.block_type[blk]);
    const int64_t prow = (int64_t)slot * M.n_pr + pcol(M, tsp, n, 0);
    cptr<VisitPack> vis = as_const(M.vadj);
    const int vb = uniform(g.vbeg), ve = uniform(g.vend);
    int cls0 = -1, cls1 = -1;
    double dprev0 = 0.0, dprev1 = 0.0;
    auto close0 = [&](int c) {
        if (c >= 0 && c <= 3) {
            const double D = row_dot<L>(T, 0, g0r, g0i);
            eacc = fma(class_weight(c), D - dprev0, eacc);
            dprev0 = D;
        }
    };
    auto close1 = [&](int c) {
        if (c >= 0 && c <= 3) {
            const double D = row_dot<L>(T, 1, g1r, g1i);
            eacc = fma(class_weight(c), D - dprev1, eacc);
            dprev1 = D;
        }
    };
    VisitPack nxt{};
    if (vb < ve) nxt = vis[vb];
    for (int vi = vb; vi < ve; ++vi) {
        const VisitPack cur = nxt;
        if (vi + 1 < ve) nxt = vis[vi + 1];
        const VDec v = decode(cur);
        const int order = v.order;
        if (order == 0) continue;  // marker of a fused row no family adds to
        const int row = (v.flags & 1) ? 1 : (v.flags & 8) ? 2 : 0;
        const int h0 = v.pslot;
        const int c = h0 >= 0 ? 4 : order - 1;
        if (row != 1 && c != cls0) {
            close0(cls0);
            cls0 = c;
        }
        if (row != 0 && c != cls1) {
            close1(cls1);
            cls1 = c;
        }
        const double H0 = h_value(M, S, prow, n, h0);
        cptr<double> cg = as_const(M.coef) + uniform(v.coef);
        if (order <= 2) {
            if (row == 0)
                fused12<L>(T, order, uniform(v.sa), H0, cg, g0r, g0i);
            else
                fused12<L>(T, order, uniform(v.sa), H0, cg, g1r, g1i);
            continue;
        }
        double tr[L + 1], ti[L + 1];
#pragma unroll
        for (int k = 0; k <= L; ++k) tr[k] = ti[k] = 0.0;
        double dd = 0.0;
        Unrolled<L, true, false, false>::run(T, v, 0, order, uniform(v.la), uniform(v.lb), uniform(v.lc), cg, cg,
                                             1.0, tr, ti, dd);
        if (row != 1) {
#pragma unroll
            for (int k = 0; k <= L; ++k) {
                g0r[k] = fma(H0, tr[k], g0r[k]);
                g0i[k] = fma(H0, ti[k], g0i[k]);
            }
        }
        if (row != 0) {
            const double H1 = row == 2 ? h_value(M, S, prow, n, uniform(v.itc)) : H0;
#pragma unroll
            for (int k = 0; k <= L; ++k) {
                g1r[k] = fma(H1, tr[k], g1r[k]);
                g1i[k] = fma(H1, ti[k], g1i[k]);
            }
        }
    }
    close0(cls0);
    close1(cls1);
    if (uniform(g.nparts) == 1) {
        // a fused row (L <= q12_fuse) is written, any other added to
        const bool first = L <= M.q12_fuse;
        double* G0 = g_row_base(M, S, blk, n, T.lane);
        double* G1 = G0 + static_cast<size_t>(M.n_lmh) * 2 * kBlock;
#pragma unroll
        for (int k = 0; k <= L; ++k) {
            const int lm = L * (L + 1) / 2 + L - k;  // lmh(L, -k)
            double* a0 = G0 + (2 * lm) * kBlock;
            double* b0 = G0 + (2 * lm + 1) * kBlock;
            double* a1 = G1 + (2 * lm) * kBlock;
            double* b1 = G1 + (2 * lm + 1) * kBlock;
            if (first) {
                *a0 = g0r[k];
                *b0 = g0i[k];
                *a1 = g1r[k];
                *b1 = g1i[k];
            } else {
                *a0 += g0r[k];
                *b0 += g0i[k];
                *a1 += g1r[k];
                *b1 += g1i[k];
            }
        }
    } else {
        // partial rows: row 0, then row 1
        double* pr = red + uniform(g.red);
        double* pr1 = pr + (L + 1) * 2 * kBlock;
#pragma unroll
        for (int k = 0; k <= L; ++k) {
            pr[(2 * k) * kBlock + T.lane] = g0r[k];
            pr[(2 * k + 1) * kBlock + T.lane] = g0i[k];
            pr1[(2 * k) * kBlock + T.lane] = g1r[k];
            pr1[(2 * k + 1) * kBlock + T.lane] = g1i[k];
        }
    }
}

template <int L, bool FEAT, bool DYN>
__device__ __forceinline__ void q34_group(const Tile& T, const DevModel& M, const DevSystem& S, const GroupQ34& g,
                                          int blk, int n, double* red, double& eacc) {
    double gr[L + 1], gi[L + 1];
    const int Sx = uniform(g.S);
#pragma unroll
    for (int k = 0; k <= L; ++k) gr[k] = gi[k] = 0.0;
    const int slot = blk * kBlock + T.lane;
    // polynomial features and H of this slot: compact row of its species,
    // column pcol(n, rank) = prow + rank
    const int tsp = uniform(S.block_type[blk]);
    const int64_t prow = (int64_t)slot * M.n_pr + pcol(M, tsp, n, 0);
    cptr<VisitPack> vis = as_const(FEAT ? M.vfeat : M.vadj);
    const int vb = uniform(g.vbeg), ve = uniform(g.vend);
    // energy classes of the visits (sorted in this order within a row):
    // linear order 1, 2, 3, 4 (class q - 1), 4 polynomial features; a linear
    // class's energy is 1/q of the row's dot product gained across it
    int cls_cur = -1;
    double dprev = 0.0;
    auto close_class = [&](int c) {
        if (c >= 0 && c <= 3) {
            const double D = row_dot<L>(T, Sx, gr, gi);
            eacc = fma(c == 0 ? 1.0 : c == 1 ? 0.5 : c == 2 ? (1.0 / 3.0) : 0.25, D - dprev, eacc);
            dprev = D;
        }
    };
    // descriptors one visit ahead: the next visit's load overlaps this one
    VisitPack nxt{};
    if (vb < ve) nxt = vis[vb];
    for (int vi = vb; vi < ve; ++vi) {
        const VisitPack cur = nxt;
        if (vi + 1 < ve) nxt = vis[vi + 1];
        const VDec v = decode(cur);
        const int order = v.order, flags = v.flags, pslot = v.pslot;
        const bool fast = (flags & 4) != 0;
        if (fast == DYN) continue;  // the other kernel's visit
        if (order == 0) continue;   // marker of a fused row no family adds to
        double d = 0.0;
        if (!FEAT) {
            const int c = pslot >= 0 ? 4 : order - 1;
            if (c != cls_cur) {
                close_class(cls_cur);
                cls_cur = c;
            }
        }
        if (!FEAT && order <= 2) {
            // fused order-1/2 adjoint: G_{S,L,-k} += H C_k a_{s_a,L,+k}
            // (order 1: G_{S,0,0} += H C)
            const double H = h_value(M, S, prow, n, pslot);
            cptr<double> c1 = as_const(M.coef) + uniform(v.coef);
            if (order == 1) {
                gr[0] = fma(H, c1[0], gr[0]);
            } else {
                const int sa = uniform(v.sa);
#pragma unroll
                for (int k = 0; k <= L; ++k) {
                    double yr, yi;
                    T.get(sa, L, k, yr, yi);
                    const double hc = H * c1[k];
                    gr[k] = fma(hc, yr, gr[k]);
                    gi[k] = fma(hc, yi, gi[k]);
                }
            }
            continue;
        }
        const int la = uniform(v.la), lb = uniform(v.lb), lc = uniform(v.lc);
        cptr<double> cg = as_const(M.coef) + uniform(v.coef);
        cptr<double> cd = as_const(M.coef) + uniform(v.dcoef);
        if (FEAT) {
            if constexpr (DYN) {
                if (order == 3)
                    visit3<L, false, true>(T, M, v, Sx, 0.0, gr, gi, d);
                else
                    visit4<L, false, true>(T, M, v, Sx, 0.0, gr, gi, d);
            } else {
                Unrolled<L, false, true, true>::run(T, v, Sx, order, la, lb, lc, cg, cd, 0.0, gr, gi, d);
            }
            S.D[prow + pslot] = d;
            continue;
        }
        const double H = h_value(M, S, prow, n, pslot);
        if constexpr (DYN) {
            if (order == 3)
                visit3<L, true, false>(T, M, v, Sx, H, gr, gi, d);
            else
                visit4<L, true, false>(T, M, v, Sx, H, gr, gi, d);
        } else {
            Unrolled<L, true, false, true>::run(T, v, Sx, order, la, lb, lc, cg, cd, H, gr, gi, d);
        }
    }
    if (!FEAT) {
        close_class(cls_cur);
        if (uniform(g.nparts) == 1) {
            // fused rows are this kernel's alone: the first kernel writes them
            const bool first = !DYN && L <= M.q12_fuse;
#pragma unroll
            for (int k = 0; k <= L; ++k) {
                const size_t ir = a_index(M, blk, n, Sx, lmh(L, -k), 0, T.lane);
                const size_t ii = a_index(M, blk, n, Sx, lmh(L, -k), 1, T.lane);
                if (first) {
                    S.G[ir] = gr[k];
                    S.G[ii] = gi[k];
                } else {
                    S.G[ir] += gr[k];
                    S.G[ii] += gi[k];
                }
            }
        } else {
            double* pr = red + uniform(g.red);
#pragma unroll
            for (int k = 0; k <= L; ++k) {
                pr[(2 * k) * kBlock + T.lane] = gr[k];
                pr[(2 * k + 1) * kBlock + T.lane] = gi[k];
            }
        }
    }
}

// A fused order-1/2 row above the LDS tile (lmax3 < L <= q12_fuse_hi): its
// visits are order 1/2 only and read a_{s_a,L,+k} from HBM (each a 64-lane
// coalesced line); energies by the same Euler classes, G written like the
// other fused rows (first write, or a partial row of a split row).  One
// instance serves every such L (entries k <= L of LM-long arrays), which
// keeps the kernel's code, and the scheduler's register choices for the
// other rows, as they are.
template <int LM>
__device__ __forceinline__ void q34_group_glob(const Tile& T, const DevModel& M, const DevSystem& S,
                                               const GroupQ34& g, int blk, int n, double* red, double& eacc) {
    const int L = uniform(g.L);
    double gr[LM + 1], gi[LM + 1];
    const int Sx = uniform(g.S);
#pragma unroll
    for (int k = 0; k <= LM; ++k) gr[k] = gi[k] = 0.0;
    const int slot = blk * kBlock + T.lane;
    const int tsp = uniform(S.block_type[blk]);
    const int64_t prow = (int64_t)slot * M.n_pr + pcol(M, tsp, n, 0);
    cptr<VisitPack> vis = as_const(M.vadj);
    const int vb = uniform(g.vbeg), ve = uniform(g.vend);
    int cls_cur = -1;
    double dprev = 0.0;
    auto close_class = [&](int c) {
        if (c >= 0 && c <= 3) {
            const double* rg = T.rowg(Sx, L);
            double D = 0.0;
#pragma unroll
            for (int k = 0; k <= LM; ++k) {
                if (k <= L) {
                    double ar, ai;
                    Tile::entryg(rg, L, -k, ar, ai);
                    D = fma(k == 0 ? 1.0 : 2.0, gr[k] * ar - gi[k] * ai, D);
                }
            }
            eacc = fma(class_weight(c), D - dprev, eacc);
            dprev = D;
        }
    };
    for (int vi = vb; vi < ve; ++vi) {
        const VDec v = decode(vis[vi]);
        const int order = v.order;
        if (order == 0 || order > 2) continue;
        const int h = v.pslot;
        const int c = h >= 0 ? 4 : order - 1;
        if (c != cls_cur) {
            close_class(cls_cur);
            cls_cur = c;
        }
        const double H = h_value(M, S, prow, n, h);
        cptr<double> c1 = as_const(M.coef) + uniform(v.coef);
        if (order == 1) {
            gr[0] = fma(H, c1[0], gr[0]);
            continue;
        }
        const double* rg = T.rowg(uniform(v.sa), L);
#pragma unroll
        for (int k = 0; k <= LM; ++k) {
            if (k <= L) {
                double yr, yi;
                Tile::entryg(rg, L, k, yr, yi);
                const double hc = H * c1[k];
                gr[k] = fma(hc, yr, gr[k]);
                gi[k] = fma(hc, yi, gi[k]);
            }
        }
    }
    close_class(cls_cur);
    if (uniform(g.nparts) == 1) {
        double* Gs = g_row_base(M, S, blk, n, T.lane) + static_cast<size_t>(Sx) * M.n_lmh * 2 * kBlock +
                     static_cast<size_t>(L * (L + 1) / 2) * 2 * kBlock;
#pragma unroll
        for (int k = 0; k <= LM; ++k) {
            if (k <= L) {
                Gs[(2 * (L - k)) * kBlock] = gr[k];  // lmh(L, -k) = L (L + 1) / 2 + L - k
                Gs[(2 * (L - k) + 1) * kBlock] = gi[k];
            }
        }
    } else {
        double* pr = red + uniform(g.red);
#pragma unroll
        for (int k = 0; k <= LM; ++k) {
            if (k <= L) {
                pr[(2 * k) * kBlock + T.lane] = gr[k];
                pr[(2 * k + 1) * kBlock + T.lane] = gi[k];
            }
        }
    }
}

// Sum of the partial rows of a split (S, L) row, in part order, into G.
__device__ __forceinline__ void q34_combine(const DevModel& M, const DevSystem& S, const GroupQ34& g, int blk, int n,
                                            int lane, const double* red, bool dyn) {
    const int L = uniform(g.L), Sx = uniform(g.S), np = uniform(g.nparts);
    const bool first = !dyn && q12_fused(M, L);  // a fused row: written, not added to
    // a species-paired group keeps rows 0 and 1 one after the other per part
    const int nrows = Sx == kPairedRows ? 2 : 1;
    const int rowlen = (L + 1) * 2 * kBlock;
    const int stride = rowlen * nrows;
    for (int r = 0; r < nrows; ++r) {
        const double* base = red + uniform(g.red) + r * rowlen;
        const int So = nrows == 2 ? r : Sx;
        for (int k = 0; k <= L; ++k) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < np; ++j) {
                sr += base[j * stride + (2 * k) * kBlock + lane];
                si += base[j * stride + (2 * k + 1) * kBlock + lane];
            }
            double& gre = S.G[a_index(M, blk, n, So, lmh(L, -k), 0, lane)];
            double& gim = S.G[a_index(M, blk, n, So, lmh(L, -k), 1, lane)];
            if (first) {
                gre = sr;
                gim = si;
            } else {
                gre += sr;
                gim += si;
            }
        }
    }
}

// largest output L with unrolled signatures (groups above it hold only
// general-loop visits)
constexpr int kMaxLUnrolled = MLIP_CAP3;

template <bool FEAT, bool DYN>
__global__ __launch_bounds__(kQ34Threads) void k_q34(DevModel M, DevSystem S) {
    extern __shared__ __attribute__((aligned(16))) double As[];
    const int blk = blockIdx.x, n = blockIdx.y;
    const int t = S.block_type[blk];
    const int per_s = M.n_lmh3 * kBlock;  // entries (re, im) per species
    double2* As2 = reinterpret_cast<double2*>(As);
    for (int s = 0; s < M.n_type; ++s) {
        // global [lm][re,im][lane] -> LDS [lm][lane](re, im)
        const double* src = S.A + a_index(M, blk, n, s, 0, 0, 0);
        for (int k = threadIdx.x; k < per_s; k += kQ34Threads) {
            const int lm = k / kBlock, ln = k % kBlock;
            As2[s * per_s + k] = double2{src[2 * lm * kBlock + ln], src[(2 * lm + 1) * kBlock + ln]};
        }
    }
    double* red = As + 2 * M.n_type * per_s;  // partial rows of split groups
    __syncthreads();
    const int wave = uniform(threadIdx.x >> 6);
    Tile T{As2, M.n_lmh3, static_cast<int>(threadIdx.x & 63), S.A + a_index(M, blk, n, 0, 0, 0, 0), M.n_lmh};
    cptr<GroupQ34> groups = as_const(FEAT ? M.gfeat : M.gadj);
    const int gb = FEAT ? M.gfeat_beg[t] : M.gadj_beg[t];
    const int ge = FEAT ? M.gfeat_end[t] : M.gadj_end[t];
    double eacc = 0.0;
    for (int gi = gb; gi < ge; ++gi) {
        const GroupQ34 g = groups[gi];
        if (uniform(g.wave) != wave) continue;
        const int L = uniform(g.L);
        // above the unrolled signatures the first kernel has only fused
        // order-1/2 rows (adjoint)
        if (!DYN && L > kMaxLUnrolled && (FEAT || L <= M.lmax3)) continue;
        if (!FEAT && L > M.lmax3) {
            // fused order-1/2 row above the tile: the first kernel's
            if constexpr (!FEAT && !DYN) q34_group_glob<kMaxLQ12>(T, M, S, g, blk, n, red, eacc);
            continue;
        }
        if (!FEAT && uniform(g.S) == kPairedRows) {
            // species-paired rows hold unrolled visits only: this kernel's
            if constexpr (!FEAT && !DYN) {
                switch (L) {
                    case 0: q34_group_pair<0>(T, M, S, g, blk, n, red, eacc); break;
                    case 1: if constexpr (kPairMaxL >= 1) q34_group_pair<1>(T, M, S, g, blk, n, red, eacc); break;
                    case 2: if constexpr (kPairMaxL >= 2) q34_group_pair<2>(T, M, S, g, blk, n, red, eacc); break;
                    case 3: if constexpr (kPairMaxL >= 3) q34_group_pair<3>(T, M, S, g, blk, n, red, eacc); break;
                    case 4: if constexpr (kPairMaxL >= 4) q34_group_pair<4>(T, M, S, g, blk, n, red, eacc); break;
                    case 5: if constexpr (kPairMaxL >= 5) q34_group_pair<5>(T, M, S, g, blk, n, red, eacc); break;
                    case 6: if constexpr (kPairMaxL >= 6) q34_group_pair<6>(T, M, S, g, blk, n, red, eacc); break;
                    case 7: if constexpr (kPairMaxL >= 7) q34_group_pair<7>(T, M, S, g, blk, n, red, eacc); break;
                    case 8: if constexpr (kPairMaxL >= 8) q34_group_pair<8>(T, M, S, g, blk, n, red, eacc); break;
                    default: break;
                }
            }
            continue;
        }
        switch (L) {
            case 0: q34_group<0, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 1: q34_group<1, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 2: q34_group<2, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 3: q34_group<3, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 4: q34_group<4, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 5: q34_group<5, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 6: q34_group<6, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 7: q34_group<7, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 8: q34_group<8, FEAT, DYN>(T, M, S, g, blk, n, red, eacc); break;
            case 9:
                if constexpr (DYN) q34_group<9, FEAT, DYN>(T, M, S, g, blk, n, red, eacc);
                break;
            case 10:
                if constexpr (DYN) q34_group<10, FEAT, DYN>(T, M, S, g, blk, n, red, eacc);
                break;
            default: break;
        }
    }
    if (!FEAT) {
        if (M.q34_red > 0) {
            __syncthreads();
            for (int gi = gb; gi < ge; ++gi) {
                const GroupQ34 g = groups[gi];
                if (uniform(g.wave) != wave || uniform(g.nparts) == 1 || uniform(g.part) != 0) continue;
                if (!DYN && uniform(g.L) > kMaxLUnrolled && uniform(g.L) <= M.lmax3) continue;
                if (DYN && uniform(g.L) > M.lmax3) continue;  // the first kernel's
                if (DYN && uniform(g.S) == kPairedRows) continue;  // the unrolled kernel's
                q34_combine(M, S, g, blk, n, T.lane, red, DYN);
            }
        }
        double* e = &S.Eq34[((size_t)(blk * kBlock + T.lane) * M.n_fn + n) * kWavesPerWg + wave];
        if (DYN)
            *e += eacc;
        else
            *e = eacc;
    }
}

template <bool FEAT, bool DYN>
hipError_t launch(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const size_t lds = sizeof(double) * (static_cast<size_t>(m.n_type) * m.n_lmh3 * 2 * kBlock + (FEAT ? 0 : m.q34_red));
    if (lds > lds_limit()) return hipErrorInvalidConfiguration;
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_q34<FEAT, DYN>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_q34<FEAT, DYN>), dim3(s.nblocks, m.n_fn), dim3(kQ34Threads), lds, st, m, s);
    return launched();
}

}  // namespace q34
}  // namespace mlip
