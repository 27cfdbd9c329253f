// Radial-only style (reference: pair_style mlip_pair,
// src/USER-MLIP/pair_mlip_pair.cpp:56-258).  Per centre i of type t:
//   d_{tc(t,s),n} = sum_{j of type s, r < r_c} g_n(r_ij)
//   E_i = sum_a w_a d_a + sum_terms w prod d,   H_a = dE_i/dd_a
// and per directed pair dE_i/d(del) = sum_n g'_n(r) H_{tc,n} del / r.  The
// pair results go through the same gather and reductions as the gtinv path.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "device_api.h"
#include "device_common.h"

namespace mlip {

namespace {

using namespace dev;

constexpr double kPiRef = 3.141592653589793;  // mlip_basis_function.cpp:26
constexpr int kPairThreads = 64;

__device__ __forceinline__ bool pair_geom(const DevModel& M, const DevSystem& S, int atom, int64_t p, double& dx,
                                          double& dy, double& dz, double& r) {
    const int nb = S.pnbr[p];
    const int c = S.pcen ? S.pcen[p] : atom;
    dx = S.x[3 * c] - S.x[3 * nb];
    dy = S.x[3 * c + 1] - S.x[3 * nb + 1];
    dz = S.x[3 * c + 2] - S.x[3 * nb + 2];
    if (S.pshift) {
        dx += S.pshift[3 * p];
        dy += S.pshift[3 * p + 1];
        dz += S.pshift[3 * p + 2];
    }
    r = sqrt(dx * dx + dy * dy + dz * dz);
    return r < M.cutoff && r > 0.0;
}

// thread per slot; NF = radial functions rounded up (accumulators by the
// neighbor's species)
template <int NF>
__global__ __launch_bounds__(kPairThreads) void k_pair_feat(DevModel M, DevSystem S) {
    const int slot = blockIdx.x * kPairThreads + threadIdx.x;
    if (slot >= S.nblocks * kBlock) return;
    const int atom = S.atom_of_slot[slot];
    double* row = S.D + static_cast<size_t>(slot) * M.n_p;
    for (int a = 0; a < M.n_p; ++a) row[a] = 0.0;
    if (atom < 0) return;
    const int t = S.elem[atom], nfn = M.n_fn;
    double acc0[NF], acc1[NF];
#pragma unroll
    for (int n = 0; n < NF; ++n) acc0[n] = acc1[n] = 0.0;
    const double pir = kPiRef / M.cutoff;
    for (int64_t p = S.poff[atom]; p < S.pend[atom]; ++p) {
        double dx, dy, dz, r;
        if (!pair_geom(M, S, atom, p, dx, dy, dz, r)) continue;
        const int s = S.elem[S.pnbr[p]];
        const double fc = 0.5 * (cos(pir * r) + 1.0);
#pragma unroll
        for (int n = 0; n < NF; ++n) {
            if (n < nfn) {
                const double d = r - M.mu[n];
                const double g = exp(-M.beta[n] * d * d) * fc;
                if (s == 0)
                    acc0[n] += g;
                else
                    acc1[n] += g;
            }
        }
    }
    for (int s = 0; s < M.n_type; ++s) {
        const int base = M.pair_tc[t][s] * nfn;
#pragma unroll
        for (int n = 0; n < NF; ++n)
            if (n < nfn) row[base + n] = s == 0 ? acc0[n] : acc1[n];
    }
}

__global__ __launch_bounds__(kPairThreads) void k_pair_poly(DevModel M, DevSystem S) {
    const int slot = blockIdx.x * kPairThreads + threadIdx.x;
    if (slot >= S.nblocks * kBlock) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) {
        S.Ep[slot] = 0.0;
        return;
    }
    const double* d = S.D + static_cast<size_t>(slot) * M.n_p;
    double* h = S.H + static_cast<size_t>(slot) * M.n_p;
    cptr<int32_t> off = as_const(M.ph_off);
    cptr<PairHEntry> ent = as_const(M.ph_ent);
    cptr<double> w = as_const(M.pair_w);
    double e = 0.0;
    for (int a = 0; a < M.n_p; ++a) {
        double hs = 0.0, he = 0.0;
        for (int k = off[a]; k < off[a + 1]; ++k) {
            const PairHEntry q = ent[k];
            double v = (q.b >= 0 ? d[q.b] : 1.0) * (q.c >= 0 ? d[q.c] : 1.0);
            hs = fma(q.w, v, hs);
            he = fma(q.ew, v, he);
        }
        h[a] = w[a] + hs;
        // the same derivative with each term weighted by 1 / its degree: the
        // reference's per-pair energy split (pair_mlip_pair.cpp:159-162)
        S.Hq[static_cast<size_t>(slot) * M.n_p + a] = w[a] + he;
        e = fma(d[a], w[a] + he, e);
    }
    S.Ep[slot] = e;
}

__global__ __launch_bounds__(kPairThreads) void k_pair_force(DevModel M, DevSystem S) {
    const int slot = blockIdx.x * kPairThreads + threadIdx.x;
    if (slot >= S.nblocks * kBlock) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) return;
    const int t = S.elem[atom], nfn = M.n_fn;
    const double* h = S.H + static_cast<size_t>(slot) * M.n_p;
    const double pir = kPiRef / M.cutoff;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int64_t p = S.poff[atom]; p < S.pend[atom]; ++p) {
        double dx, dy, dz, r;
        double gx = 0.0, gy = 0.0, gz = 0.0;
        if (pair_geom(M, S, atom, p, dx, dy, dz, r)) {
            const int s = S.elem[S.pnbr[p]];
            const double* hr = h + M.pair_tc[t][s] * nfn;
            double sa, ca;
            sincos(pir * r, &sa, &ca);
            const double fc = 0.5 * (ca + 1.0), fcd = -0.5 * pir * sa;
            double de = 0.0;
            for (int n = 0; n < nfn; ++n) {
                const double d = r - M.mu[n];
                const double b = exp(-M.beta[n] * d * d);
                de = fma(b * (fcd - 2.0 * M.beta[n] * d * fc), hr[n], de);
            }
            const double ir = de / r;
            gx = ir * dx;
            gy = ir * dy;
            gz = ir * dz;
        }
        S.gpair[3 * p] = gx;
        S.gpair[3 * p + 1] = gy;
        S.gpair[3 * p + 2] = gz;
        v0 -= dx * gx;
        v1 -= dy * gy;
        v2 -= dz * gz;
        v3 -= dx * gy;
        v4 -= dx * gz;
        v5 -= dy * gz;
    }
    double* vo = S.vown + 6 * static_cast<size_t>(atom);
    vo[0] = v0;
    vo[1] = v1;
    vo[2] = v2;
    vo[3] = v3;
    vo[4] = v4;
    vo[5] = v5;
}

// Per directed pair: energy share e_p = sum_n g_n(r) He_{tc,n} (He: the
// split derivative k_pair_poly leaves in Hq) and the pair's virial
// del (x) (-g_p); the reference's half-pair evdwl is e_p + e_p' over the two
// directions (pair_mlip_pair.cpp:196-212).
__global__ __launch_bounds__(kPairThreads) void k_pair_model_energy(DevModel M, DevSystem S, double* epv) {
    const int slot = blockIdx.x * kPairThreads + threadIdx.x;
    if (slot >= S.nblocks * kBlock) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) return;
    const int t = S.elem[atom], nfn = M.n_fn;
    const double* he = S.Hq + static_cast<size_t>(slot) * M.n_p;
    const double pir = kPiRef / M.cutoff;
    for (int64_t p = S.poff[atom]; p < S.pend[atom]; ++p) {
        double dx, dy, dz, r;
        double e = 0.0, v[6] = {0, 0, 0, 0, 0, 0};
        if (pair_geom(M, S, atom, p, dx, dy, dz, r)) {
            const int s = S.elem[S.pnbr[p]];
            const double* hr = he + M.pair_tc[t][s] * nfn;
            const double fc = 0.5 * (cos(pir * r) + 1.0);
            for (int n = 0; n < nfn; ++n) {
                const double d = r - M.mu[n];
                e = fma(exp(-M.beta[n] * d * d) * fc, hr[n], e);
            }
            const double gx = S.gpair[3 * p], gy = S.gpair[3 * p + 1], gz = S.gpair[3 * p + 2];
            v[0] = -dx * gx;
            v[1] = -dy * gy;
            v[2] = -dz * gz;
            v[3] = -dx * gy;
            v[4] = -dx * gz;
            v[5] = -dy * gz;
        }
        epv[7 * p] = e;
        for (int c = 0; c < 6; ++c) epv[7 * p + 1 + c] = v[c];
    }
}

}  // namespace

hipError_t launch_pair_model_energy(const DevModel& m, const DevSystem& s, double* epv, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_model_energy, dim3((nslot + kPairThreads - 1) / kPairThreads), dim3(kPairThreads), 0,
                       st, m, s, epv);
    return launched();
}

hipError_t launch_pair_features(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot == 0) return hipSuccess;
    const dim3 grid((nslot + kPairThreads - 1) / kPairThreads), block(kPairThreads);
    if (m.n_fn <= 8)
        hipLaunchKernelGGL(k_pair_feat<8>, grid, block, 0, st, m, s);
    else if (m.n_fn <= 16)
        hipLaunchKernelGGL(k_pair_feat<16>, grid, block, 0, st, m, s);
    else
        hipLaunchKernelGGL(k_pair_feat<32>, grid, block, 0, st, m, s);
    return launched();
}

hipError_t launch_pair_poly(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_poly, dim3((nslot + kPairThreads - 1) / kPairThreads), dim3(kPairThreads), 0, st, m, s);
    return launched();
}

hipError_t launch_pair_force(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_force, dim3((nslot + kPairThreads - 1) / kPairThreads), dim3(kPairThreads), 0, st, m,
                       s);
    return launched();
}

}  // namespace mlip
