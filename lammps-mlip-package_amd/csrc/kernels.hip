// HIP kernels of the gtinv energy/force path for CDNA4 (gfx950).
//
// Pipeline for one evaluation (reference: PairMLIPGtinv::compute,
// src/USER-MLIP/pair_mlip_gtinv.cpp:63-125, and docs/formulation.md):
//   k_anlm   a_{i,s,n,lm} = sum_j g_n(r_ij) Y_lm(r_ij)          (:349-447)
//   k_q12    order-1/2 invariants: polynomial features or adjoint G
//   k_q34    order-3/4 invariants: polynomial features or adjoint G,
//            a_nlm tile of 64 same-species centres resident in LDS
//   rocBLAS  Hq = D W (quadratic form of the polynomial model)
//   k_poly   E_i polynomial part and H = dE_i/dd                 (:321-343,481-521)
//   k_force  per-pair dE_i/d(del) from G and dphi/d(del)          (:161-256)
//   k_gather atomics-free per-position force sums, deterministic
//   k_energy per-centre energies and virial, fixed-order reductions
//
// Conventions shared with the reference CPU style: del = x_centre - x_neighbor,
// a_nlm uses Y_lm(del) without conjugation, pairs with r < cutoff (strict).
// Only m <= 0 entries of a and G are stored (index l(l+1)/2 + l + m); the
// m > 0 entries follow from a_{l,m} = (-1)^m conj(a_{l,-m}).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "device_api.h"
#include "device_common.h"

namespace mlip {

namespace {

using namespace dev;

constexpr double kPiRef = 3.141592653589793;  // mlip_basis_function.cpp:26
constexpr int kL1 = kMaxL + 1;

// Normalised associated-Legendre recursion without the sin^m factor:
//   Qb_m^m      = c_diag[m]
//   Qb_{m+1}^m  = c_sub[m] z Qb_m^m
//   Qb_l^m      = c_alm[l][m] (z Qb_{l-1}^m - c_ainv[l][m] Qb_{l-2}^m)
// Y_l^m = Qb_l^m u^m, u = (x + i y)/r, Condon-Shortley phase included.
__constant__ double c_diag[kL1 + 1];
__constant__ double c_sub[kL1];
__constant__ double c_alm[kL1][kL1];
__constant__ double c_ainv[kL1][kL1];
__constant__ double c_grad[kL1][kL1];  // sqrt((l+m+1)(l-m))

struct PairGeom {
    double dx, dy, dz, r;
    bool valid;
};

__device__ __forceinline__ PairGeom pair_geom(const DevModel& M, const DevSystem& S, int owner, int64_t p) {
    PairGeom g;
    const int nb = S.pnbr[p];
    const int c = S.pcen ? S.pcen[p] : owner;
    g.dx = S.x[3 * c] - S.x[3 * nb];
    g.dy = S.x[3 * c + 1] - S.x[3 * nb + 1];
    g.dz = S.x[3 * c + 2] - S.x[3 * nb + 2];
    if (S.pshift) {
        g.dx += S.pshift[3 * p];
        g.dy += S.pshift[3 * p + 1];
        g.dz += S.pshift[3 * p + 2];
    }
    g.r = sqrt(g.dx * g.dx + g.dy * g.dy + g.dz * g.dz);
    g.valid = (g.r < M.cutoff) && (g.r > 0.0);
    return g;
}

// ---------------------------------------------------------------- a_nlm
// One workgroup per centre.  Phase 1: four threads per neighbor of a chunk
// of 32 evaluate g_n and the Y_lm (m <= 0) columns into LDS.  Phase 2:
// thread = lm, accumulates sum_j g_n(j) Y_lm(j) for both species in registers.
template <int LMAX>
__global__ __launch_bounds__(128) void k_anlm(DevModel M, DevSystem S) {
    constexpr int NLMH = (LMAX + 1) * (LMAX + 2) / 2;
    __shared__ double2 ys[kAnlmChunk][NLMH];
    __shared__ double gs[kAnlmChunk][kMaxFn];
    __shared__ int sp[kAnlmChunk];
    const int slot = xcd_slot(blockIdx.x, S.nblocks);
    if (slot < 0) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) return;
    const int tid = threadIdx.x;
    const int nfn = M.n_fn;
    double a0r[kMaxFn], a0i[kMaxFn], a1r[kMaxFn], a1i[kMaxFn];
#pragma unroll
    for (int n = 0; n < kMaxFn; ++n) a0r[n] = a0i[n] = a1r[n] = a1i[n] = 0.0;
    const int p0 = S.poff[atom], p1 = S.poff[atom + 1];
    for (int c0 = p0; c0 < p1; c0 += kAnlmChunk) {
        {
            // four threads per neighbor: part k evaluates the radial functions
            // n = k, k+4, ... and the Y_lm columns m = k, k+4, ...
            const int j = tid >> 2, part = tid & 3;
            const int64_t p = c0 + j;
            int spec = -1;
            if (p < p1) {
                PairGeom g = pair_geom(M, S, atom, p);
                if (g.valid) {
                    spec = S.elem[S.pnbr[p]];
                    const double fc = 0.5 * (cos(kPiRef * g.r / M.cutoff) + 1.0);
#pragma unroll 1
                    for (int n = part; n < nfn; n += 4) {
                        const double d = g.r - M.mu[n];
                        gs[j][n] = exp(-M.beta[n] * (d * d)) * fc;
                    }
                    const double ir = 1.0 / g.r;
                    const double x = g.dx * ir, y = g.dy * ir, z = g.dz * ir;
                    // conj(u)^part and conj(u)^4
                    double vr = 1.0, vi = 0.0;
                    for (int k = 0; k < part; ++k) {
                        const double nr = vr * x + vi * y, ni = vi * x - vr * y;
                        vr = nr;
                        vi = ni;
                    }
                    const double u2r = x * x - y * y, u2i = -2.0 * x * y;
                    const double u4r = u2r * u2r - u2i * u2i, u4i = 2.0 * u2r * u2i;
                    // stored Y_{l,-m} = (-1)^m Qb_l^m conj(u)^m
#pragma unroll 1
                    for (int m = part; m <= LMAX; m += 4) {
                        const double sg = (m & 1) ? -1.0 : 1.0;
                        double q2 = 0.0, q1 = c_diag[m];
                        ys[j][lmh(m, -m)] = make_double2(sg * q1 * vr, sg * q1 * vi);
#pragma unroll 1
                        for (int l = m + 1; l <= LMAX; ++l) {
                            const double q = (l == m + 1) ? c_sub[m] * z * q1
                                                          : c_alm[l][m] * (z * q1 - c_ainv[l][m] * q2);
                            ys[j][lmh(l, -m)] = make_double2(sg * q * vr, sg * q * vi);
                            q2 = q1;
                            q1 = q;
                        }
                        const double nr = vr * u4r - vi * u4i, ni = vr * u4i + vi * u4r;
                        vr = nr;
                        vi = ni;
                    }
                }
            }
            if (part == 0) sp[j] = spec;
        }
        __syncthreads();
        const int nc = min(kAnlmChunk, p1 - c0);
        if (tid < NLMH) {
            for (int j = 0; j < nc; ++j) {
                const int spec = sp[j];
                if (spec < 0) continue;
                const double2 y = ys[j][tid];
                if (spec == 0) {
#pragma unroll
                    for (int n = 0; n < kMaxFn; ++n)
                        if (n < nfn) {
                            const double gn = gs[j][n];
                            a0r[n] = fma(gn, y.x, a0r[n]);
                            a0i[n] = fma(gn, y.y, a0i[n]);
                        }
                } else {
#pragma unroll
                    for (int n = 0; n < kMaxFn; ++n)
                        if (n < nfn) {
                            const double gn = gs[j][n];
                            a1r[n] = fma(gn, y.x, a1r[n]);
                            a1i[n] = fma(gn, y.y, a1i[n]);
                        }
                }
            }
        }
        __syncthreads();
    }
    if (tid < NLMH) {
        const int blk = slot / kBlock, lane = slot % kBlock;
#pragma unroll
        for (int n = 0; n < kMaxFn; ++n)
            if (n < nfn) {
                S.A[a_index(M, blk, n, 0, tid, 0, lane)] = a0r[n];
                S.A[a_index(M, blk, n, 0, tid, 1, lane)] = a0i[n];
                if (M.n_type > 1) {
                    S.A[a_index(M, blk, n, 1, tid, 0, lane)] = a1r[n];
                    S.A[a_index(M, blk, n, 1, tid, 1, lane)] = a1i[n];
                }
            }
    }
}

// ---------------------------------------------------------------- order 1/2
// Wave = 64 slots of one block at one radial index n; a_{s,L,m} for the
// current L is held in registers, G_{s,L,m} accumulated in registers.
template <int L, bool FEAT>
__device__ void q12_level(const DevModel& M, const DevSystem& S, int blk, int n, int lane, int t,
                          double& eacc) {
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
        if (!FEAT) H = pslot >= 0 ? S.H[(size_t)slot * M.n_p + n * M.n_pitc + pslot] : M.weight[n * M.n_gtinv + v.itc];
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
            S.D[(size_t)slot * M.n_p + n * M.n_pitc + pslot] = d;
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
                               double& e) {
        if (L <= M.maxl) q12_level<L, FEAT>(M, S, blk, n, lane, t, e);
        Q12Levels<L + 1, LMAX, FEAT>::run(M, S, blk, n, lane, t, e);
    }
};
template <int LMAX, bool FEAT>
struct Q12Levels<LMAX + 1, LMAX, FEAT> {
    __device__ static void run(const DevModel&, const DevSystem&, int, int, int, int, double&) {}
};

template <bool FEAT>
__global__ __launch_bounds__(64) void k_q12(DevModel M, DevSystem S) {
    const int blk = blockIdx.x, n = blockIdx.y, lane = threadIdx.x;
    const int t = S.block_type[blk];
    double e = 0.0;
    if (FEAT) {
        // features of families the centre's species does not see are zero
        // (the reference never accumulates them); the visits below and the
        // order-3/4 feature kernel overwrite the active ones
        double* row = S.D + (size_t)(blk * kBlock + lane) * M.n_p + n * M.n_pitc;
        for (int r = 0; r < M.n_pitc; ++r) row[r] = 0.0;
    }
    Q12Levels<0, kMaxL, FEAT>::run(M, S, blk, n, lane, t, e);
    if (!FEAT) S.Eq12[(size_t)(blk * kBlock + lane) * M.n_fn + n] = e;
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
        const double* d = S.D + (size_t)slot * M.n_p;
        const double* hq = S.Hq + (size_t)slot * M.n_p;
        double* h = S.H + (size_t)slot * M.n_p;
        for (int a = lane; a < M.n_p; a += 64) {
            double h3 = 0.0;
            for (int k = M.c3_off[a]; k < M.c3_off[a + 1]; ++k)
                h3 = fma(M.c3_w[k], d[M.c3_bc[2 * k]] * d[M.c3_bc[2 * k + 1]], h3);
            const double la = M.lin_p[a];
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

// ---------------------------------------------------------------- forces
// One workgroup per centre; G of the centre (both species) staged in LDS as
// [s][lm][n][re,im].  Thread = neighbor pair; Y_lm and grad Y_lm by the same
// recursion as k_anlm, contracted with
//   Gp_lm = sum_n g'_n G_nlm,  Gpp_lm = sum_n g_n G_nlm
//   dE/d(del) = sum_lm w_m Re(Gp Y) rhat + w_m Re(Gpp . grad Y),  w_0 = 1, w_{m<0} = 2
template <int LMAX, int NT>
__global__ __launch_bounds__(NT) void k_force(DevModel M, DevSystem S) {
    constexpr int NLMH = (LMAX + 1) * (LMAX + 2) / 2;
    extern __shared__ double Gs[];
    __shared__ double vred[6][NT];
    const int slot = xcd_slot(blockIdx.x, S.nblocks);
    if (slot < 0) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) return;
    const int nfn = M.n_fn, ns = M.n_type;
    const int blk = slot / kBlock, lane = slot % kBlock;
    for (int k = threadIdx.x; k < ns * NLMH * nfn; k += blockDim.x) {
        const int n = k % nfn;
        const int lm = (k / nfn) % NLMH;
        const int s = k / (nfn * NLMH);
        Gs[2 * k] = S.G[a_index(M, blk, n, s, lm, 0, lane)];
        Gs[2 * k + 1] = S.G[a_index(M, blk, n, s, lm, 1, lane)];
    }
    __syncthreads();
    const int p0 = S.poff[atom], p1 = S.poff[atom + 1];
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        PairGeom pg = pair_geom(M, S, atom, p);
        double fx = 0.0, fy = 0.0, fz = 0.0;
        if (pg.valid) {
            const int spec = S.elem[S.pnbr[p]];
            const double* Gsp = Gs + (size_t)spec * NLMH * nfn * 2;
            const double ir = 1.0 / pg.r;
            const double x = pg.dx * ir, y = pg.dy * ir, z = pg.dz * ir;
            const double arg = kPiRef * pg.r / M.cutoff;
            const double fc = 0.5 * (cos(arg) + 1.0);
            const double fcd = -0.5 * kPiRef / M.cutoff * sin(arg);
            double gn[kMaxFn], gd[kMaxFn];
#pragma unroll
            for (int n = 0; n < kMaxFn; ++n) {
                gn[n] = gd[n] = 0.0;
                if (n < nfn) {
                    const double dd = pg.r - M.mu[n];
                    const double b = exp(-M.beta[n] * (dd * dd));
                    const double bd = -2.0 * M.beta[n] * dd * b;
                    gn[n] = b * fc;
                    gd[n] = bd * fc + b * fcd;
                }
            }
            double er = 0.0;      // sum w Re(Gp Y)
            double tx = 0.0, ty = 0.0, tz = 0.0;  // sum w Re(Gpp grad Y) * r
            const double zz = 1.0 - z * z;
            // Columns m and m+1 of the recursion advance in lockstep along l,
            // each carried in two scalars, so no per-lane array is indexed.
            double cr = 1.0, ci = 0.0;  // conj(u)^m
            double mr = 0.0, mi = 0.0;  // conj(u)^(m-1), used for m >= 1
#pragma unroll 1
            for (int m = 0; m <= LMAX; ++m) {
                const double sg = (m & 1) ? -1.0 : 1.0;
                const double w = (m == 0) ? 1.0 : 2.0;
                const double pr = cr * x + ci * y, pi = ci * x - cr * y;  // conj(u)^(m+1)
                double a1 = 0.0, a2 = 0.0;  // Qb_{l-1}^m, Qb_{l-2}^m
                double b1 = 0.0, b2 = 0.0;  // Qb_{l-1}^{m+1}, Qb_{l-2}^{m+1}
#pragma unroll 1
                for (int l = m; l <= LMAX; ++l) {
                    double q;
                    if (l == m)
                        q = c_diag[m];
                    else if (l == m + 1)
                        q = c_sub[m] * z * a1;
                    else
                        q = c_alm[l][m] * (z * a1 - c_ainv[l][m] * a2);
                    a2 = a1;
                    a1 = q;
                    double qp = 0.0;  // Qb_l^{m+1}
                    if (l == m + 1)
                        qp = c_diag[m + 1];
                    else if (l == m + 2)
                        qp = c_sub[m + 1] * z * b1;
                    else if (l > m + 2)
                        qp = c_alm[l][m + 1] * (z * b1 - c_ainv[l][m + 1] * b2);
                    b2 = b1;
                    b1 = qp;
                    const int lm = lmh(l, -m);
                    // contract G over n
                    double gpr = 0.0, gpi = 0.0, gqr = 0.0, gqi = 0.0;
                    const double* gl = Gsp + (size_t)lm * nfn * 2;
#pragma unroll
                    for (int n = 0; n < kMaxFn; ++n)
                        if (n < nfn) {
                            const double grr = gl[2 * n], gii = gl[2 * n + 1];
                            gpr = fma(gd[n], grr, gpr);
                            gpi = fma(gd[n], gii, gpi);
                            gqr = fma(gn[n], grr, gqr);
                            gqi = fma(gn[n], gii, gqi);
                        }
                    // Y_{l,-m} = sg q conj(u)^m
                    const double yr = sg * q * cr, yi = sg * q * ci;
                    er = fma(w, gpr * yr - gpi * yi, er);
                    // grad Y_{l,-m} * r = sg conj(grad Y_l^m * r)
                    double gxr, gxi, gyr, gyi, gzr, gzi;
                    if (m == 0) {
                        const double c = c_grad[l][0] * qp;  // Qb_l^1 (0 for l = 0)
                        gxr = c * z * x;
                        gxi = 0.0;
                        gyr = c * z * y;
                        gyi = 0.0;
                        gzr = -c * zz;
                        gzi = 0.0;
                    } else {
                        const double c = c_grad[l][m];
                        const double A = m * z * q + zz * c * qp;
                        const double B = m * q - z * c * qp;
                        const double E = m * (1.0 + z * z) * q + z * zz * c * qp;
                        // conj of: gx = (-B u^{m+1} + E u^{m-1})/2, gy = (i/2)(B u^{m+1} + E u^{m-1}),
                        // gz = -A u^m, with conj(u)^{m+1} = (pr, pi), conj(u)^{m-1} = (mr, mi)
                        gxr = 0.5 * (-B * pr + E * mr);
                        gxi = 0.5 * (-B * pi + E * mi);
                        const double zr_ = B * pr + E * mr, zi_ = B * pi + E * mi;
                        gyr = 0.5 * zi_;
                        gyi = -0.5 * zr_;
                        gzr = -A * cr;
                        gzi = -A * ci;
                    }
                    tx = fma(w * sg, gqr * gxr - gqi * gxi, tx);
                    ty = fma(w * sg, gqr * gyr - gqi * gyi, ty);
                    tz = fma(w * sg, gqr * gzr - gqi * gzi, tz);
                }
                // advance powers: conj(u)^m becomes the next iteration's m-1 power
                mr = cr;
                mi = ci;
                cr = pr;
                ci = pi;
            }
            // dE/d(del) = er * rhat + t / r
            fx = er * x + tx * ir;
            fy = er * y + ty * ir;
            fz = er * z + tz * ir;
        }
        S.gpair[3 * p] = fx;
        S.gpair[3 * p + 1] = fy;
        S.gpair[3 * p + 2] = fz;
        // virial: del (x) F_centre with F_centre = -dE/d(del)
        v0 -= pg.dx * fx;
        v1 -= pg.dy * fy;
        v2 -= pg.dz * fz;
        v3 -= pg.dx * fy;
        v4 -= pg.dx * fz;
        v5 -= pg.dy * fz;
    }
    vred[0][threadIdx.x] = v0;
    vred[1][threadIdx.x] = v1;
    vred[2][threadIdx.x] = v2;
    vred[3][threadIdx.x] = v3;
    vred[4][threadIdx.x] = v4;
    vred[5][threadIdx.x] = v5;
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = 0.0;
        for (int k = 0; k < NT; ++k) s += vred[threadIdx.x][k];
        S.vown[6 * atom + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------- gather
__global__ __launch_bounds__(256) void k_gather(DevSystem S) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S.nall) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int e = S.goff[k]; e < S.goff[k + 1]; ++e) {
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
    if (S.atom_of_slot[slot] >= 0) {
        e = M.n_p > 0 ? S.Ep[slot] : 0.0;
        for (int n = 0; n < M.n_fn; ++n) {
            e += S.Eq12[(size_t)slot * M.n_fn + n];
            if (M.has_q34)
                for (int w = 0; w < kWavesPerWg; ++w) e += S.Eq34[((size_t)slot * M.n_fn + n) * kWavesPerWg + w];
        }
    }
    S.e_slot[slot] = e;
}

// Per-structure energy and virial of a batch: one workgroup per structure,
// fixed-order tree over its (contiguous) atoms.
__global__ __launch_bounds__(256) void k_struct_reduce(DevSystem S, const int32_t* soff, double* out) {
    __shared__ double sh[256];
    const int st = blockIdx.x;
    const int a0 = soff[st], a1 = soff[st + 1];
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

__global__ __launch_bounds__(256) void k_copy(double4* __restrict__ dst, const double4* __restrict__ src, int64_t n) {
    for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "mlip_hip: %s failed: %s\n", what, hipGetErrorString(e));
    }
}

}  // namespace

void init_device_constants() {
    double diag[kL1 + 1], sub[kL1], alm[kL1][kL1], ainv[kL1][kL1], grad[kL1][kL1];
    long double dg = 1.0L / sqrtl(4.0L * 3.14159265358979323846264338327950288L);
    for (int m = 0; m <= kL1; ++m) {
        if (m > 0) dg *= -sqrtl((2.0L * m + 1.0L) / (2.0L * m));
        diag[m] = static_cast<double>(dg);
    }
    for (int m = 0; m < kL1; ++m) sub[m] = static_cast<double>(sqrtl(2.0L * m + 3.0L));
    for (int l = 0; l < kL1; ++l)
        for (int m = 0; m < kL1; ++m) {
            alm[l][m] = ainv[l][m] = grad[l][m] = 0.0;
            if (l >= m + 2) {
                const long double a = sqrtl((4.0L * l * l - 1.0L) / ((long double)l * l - (long double)m * m));
                const long double a1 =
                    sqrtl((4.0L * (l - 1) * (l - 1) - 1.0L) / ((long double)(l - 1) * (l - 1) - (long double)m * m));
                alm[l][m] = static_cast<double>(a);
                ainv[l][m] = static_cast<double>(1.0L / a1);
            }
            if (m <= l) grad[l][m] = static_cast<double>(sqrtl((long double)(l + m + 1) * (l - m)));
        }
    check(hipMemcpyToSymbol(HIP_SYMBOL(c_diag), diag, sizeof diag), "copy c_diag");
    check(hipMemcpyToSymbol(HIP_SYMBOL(c_sub), sub, sizeof sub), "copy c_sub");
    check(hipMemcpyToSymbol(HIP_SYMBOL(c_alm), alm, sizeof alm), "copy c_alm");
    check(hipMemcpyToSymbol(HIP_SYMBOL(c_ainv), ainv, sizeof ainv), "copy c_ainv");
    check(hipMemcpyToSymbol(HIP_SYMBOL(c_grad), grad, sizeof grad), "copy c_grad");
}

static int slot_grid(const DevSystem& s) { return 8 * kBlock * ((s.nblocks + 7) / 8); }

#define MLIP_LMAX_CASES(X) \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

void launch_anlm(const DevModel& m, const DevSystem& s, int, hipStream_t st) {
    if (s.nblocks == 0) return;
    const int grid = slot_grid(s);
    switch (m.maxl) {
#define X(L)                                                          \
    case L:                                                           \
        hipLaunchKernelGGL(k_anlm<L>, dim3(grid), dim3(128), 0, st, m, s); \
        break;
        MLIP_LMAX_CASES(X)
#undef X
        default: break;
    }
    check(hipGetLastError(), "k_anlm launch");
}

void launch_q12(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st) {
    if (s.nblocks == 0) return;
    if (feat)
        hipLaunchKernelGGL(k_q12<true>, dim3(s.nblocks, m.n_fn), dim3(64), 0, st, m, s);
    else
        hipLaunchKernelGGL(k_q12<false>, dim3(s.nblocks, m.n_fn), dim3(64), 0, st, m, s);
    check(hipGetLastError(), "k_q12 launch");
}

void launch_poly(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot == 0 || m.n_p == 0) return;
    hipLaunchKernelGGL(k_poly, dim3((nslot + 3) / 4), dim3(256), 0, st, m, s);
    check(hipGetLastError(), "k_poly launch");
}

template <int NT>
static void launch_force_nt(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int grid = slot_grid(s);
    const size_t lds = sizeof(double) * 2 * m.n_type * m.n_lmh * m.n_fn;
    switch (m.maxl) {
#define X(L)                                                                                        \
    case L:                                                                                         \
        check(hipFuncSetAttribute(reinterpret_cast<const void*>(k_force<L, NT>),                    \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)), \
              "k_force lds attr");                                                                  \
        hipLaunchKernelGGL((k_force<L, NT>), dim3(grid), dim3(NT), lds, st, m, s);                  \
        break;
        MLIP_LMAX_CASES(X)
#undef X
        default: break;
    }
}

void launch_force(const DevModel& m, const DevSystem& s, int threads, hipStream_t st) {
    if (s.nblocks == 0) return;
    switch (threads) {
        case 64: launch_force_nt<64>(m, s, st); break;
        case 192: launch_force_nt<192>(m, s, st); break;
        case 256: launch_force_nt<256>(m, s, st); break;
        default: launch_force_nt<128>(m, s, st); break;
    }
    check(hipGetLastError(), "k_force launch");
}

void launch_gather(const DevModel&, const DevSystem& s, hipStream_t st) {
    if (s.nall == 0) return;
    hipLaunchKernelGGL(k_gather, dim3((s.nall + 255) / 256), dim3(256), 0, st, s);
    check(hipGetLastError(), "k_gather launch");
}

void launch_copy(void* dst, const void* src, int64_t nbytes, int reps, hipStream_t st) {
    const int64_t n = nbytes / static_cast<int64_t>(sizeof(double4));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, st, static_cast<double4*>(dst),
                           static_cast<const double4*>(src), n);
    check(hipGetLastError(), "k_copy launch");
}

void launch_struct_reduce(const DevSystem& s, const int32_t* soff, int nstruct, double* out, hipStream_t st) {
    if (nstruct <= 0) return;
    hipLaunchKernelGGL(k_struct_reduce, dim3(nstruct), dim3(256), 0, st, s, soff, out);
    check(hipGetLastError(), "k_struct_reduce launch");
}

void launch_energy_virial(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nslot = s.nblocks * kBlock;
    if (nslot > 0) hipLaunchKernelGGL(k_energy_slot, dim3((nslot + 255) / 256), dim3(256), 0, st, m, s);
    const int parts = 120;
    hipLaunchKernelGGL(k_reduce1, dim3(parts), dim3(kRedThreads), 0, st, s.e_slot, nslot, 1, 1, s.red);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kRedThreads), 0, st, s.red, parts, 1, s.out);
    hipLaunchKernelGGL(k_reduce1, dim3(parts), dim3(kRedThreads), 0, st, s.vown, s.nown, 6, 6, s.red + parts);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kRedThreads), 0, st, s.red + parts, parts, 6, s.out + 1);
    check(hipGetLastError(), "reduction launch");
}

}  // namespace mlip
