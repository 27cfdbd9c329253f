// Per-pair kernels of the gtinv path: a_nlm accumulation and pair forces,
// each in FP64 and in the mixed mode (FP32 per-pair geometry, radial
// functions, Y_lm and their gradients; FP64 accumulation of a_nlm, FP64
// pair forces out of the kernel, FP64 virial).
//
// Reference: PairMLIPGtinv::compute_anlm (src/USER-MLIP/pair_mlip_gtinv.cpp:349-447)
// and the force loop compute_energy_and_force_for_each_atom (:161-256);
// formulas in docs/formulation.md:126-154,221-280,351-411.
//
// Y_l^m(del) = Qb_l^m u^m with u = (x + i y)/r and the normalised Legendre
// recursion without the sin^m factor:
//   Qb_m^m = diag[m],  Qb_l^m = alm[l][m] (z Qb_{l-1}^m - ainv[l][m] Qb_{l-2}^m)
// where alm[m+1][m] = sqrt(2m+3) and Qb_{m-1}^m = 0 make the first step
// the same formula.  Only m <= 0 entries are formed: Y_{l,-m} = (-1)^m Qb conj(u)^m.
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

template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
    using type = double2;
};
template <>
struct Vec2<float> {
    using type = float2;
};

template <typename T>
__device__ __forceinline__ const YTab<T>* ytab_of(const DevModel& M);
template <>
__device__ __forceinline__ const YTab<double>* ytab_of<double>(const DevModel& M) {
    return M.yt64;
}
template <>
__device__ __forceinline__ const YTab<float>* ytab_of<float>(const DevModel& M) {
    return M.yt32;
}

// sin and cos of the cutoff argument a = pi r / r_c, 0 < r < r_c, so a lies
// in (0, pi) and needs no range reduction: t = a (a <= pi/2) or pi - a
// (exact by Sterbenz, plus the low part of pi), sin a = sin t, cos a = +-cos t,
// each a Taylor series in t^2 through t^21 and t^22 (remainder below 2e-18 for
// |t| <= pi/2; Horner's rounding, about one ulp, dominates).  The coefficients
// (-1)^k / (2k+1)! and (-1)^k / (2k)!, k >= 1, are read from LDS through a
// base the compiler cannot take as constant, so they are loaded where the
// series runs: the library routine's reduction constants, materialised once
// per thread and held across the pair loop, were spilled to scratch.
constexpr int kTrigSin = 10, kTrigCos = 11;
constexpr int kTrigN = kTrigSin + kTrigCos;
typedef const double __attribute__((address_space(3))) lds_cdouble;

__device__ __forceinline__ double trig_coef(int i) {
    const int k = i < kTrigSin ? i + 1 : i - kTrigSin + 1;
    const int n = i < kTrigSin ? 2 * k + 1 : 2 * k;
    double f = 1.0;
    for (int j = 2; j <= n; ++j) f *= static_cast<double>(j);
    return ((k & 1) ? -1.0 : 1.0) / f;
}

// q: the coefficients in LDS, as an opaque LDS pointer (lds_opaque).
__device__ __forceinline__ lds_cdouble* lds_opaque(const double* p) {
    lds_cdouble* q = (lds_cdouble*)p;
    asm volatile("" : "+v"(q));
    return q;
}

__device__ __forceinline__ void sin_cos(double a, lds_cdouble* q, double* s, double* c) {
    constexpr double kPiHi = 3.141592653589793;       // pi rounded to double
    constexpr double kPiLo = 1.2246467991473532e-16;  // pi - kPiHi
    const bool up = a > 0.5 * kPiHi;
    const double t = up ? (kPiHi - a) + kPiLo : a;
    const double u = t * t;
    double ps = q[kTrigSin - 1];
#pragma unroll
    for (int k = kTrigSin - 2; k >= 0; --k) ps = fma(ps, u, q[k]);
    double pc = q[kTrigN - 1];
#pragma unroll
    for (int k = kTrigN - 2; k >= kTrigSin; --k) pc = fma(pc, u, q[k]);
    const double st = fma(t * u, ps, t);
    const double ct = fma(u, pc, 1.0);
    *s = st;
    *c = up ? -ct : ct;
}

struct PairDel {
    double dx, dy, dz, r;
    bool valid;
};

// del = x_centre - x_neighbor (+ image shift), strict r < cutoff.  Always
// FP64: the difference of absolute positions and the cutoff test must not
// depend on the precision mode.
__device__ __forceinline__ PairDel pair_del(const DevModel& M, const DevSystem& S, int owner, int64_t p) {
    PairDel g;
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
// One workgroup of NTH threads per centre (192, or 256 when l_max > 18).
// Neighbors are taken in blocks of NTH (geometry once per neighbor into
// LDS), then in chunks of kAnlmJ: phase 1 has NTH / kAnlmJ threads per
// neighbor evaluate g_n and the Y_lm columns m = part, part + NTH / kAnlmJ,
// ... into LDS; phase 2 has thread (lm, group) accumulate a_{s,n,lm} for its
// NPT radial indices n = n0 + group*NPT + k and both species in FP64
// registers, n0 = blockIdx.y * ngroup * NPT (models with more radial
// functions than one pass holds run several passes over the neighbors).
constexpr int kAnlmJ = 16;
static_assert(kAnlmJ <= 64, "a chunk's neighbor species fit one per lane of a wave");
// Size classes of the per-workgroup tables: models with l_max <= 12 whose
// radial passes read at most 16 g_n per neighbor keep the compact tables
// (the LDS they take sets how many workgroups share a CU); larger models use
// the full l_max 20 tables and 48-wide g_n rows (every pass's NPT-wide read
// stays inside the row).
constexpr int kSmallL1 = 13, kSmallGW = 16;
constexpr int kLargeGW = kMaxFn + 16;

template <typename T, int NTH, int KL1, int GW>
struct AnlmLds {
    T diag[KL1 + 1];
    T alm[KL1][KL1];
    T ainv[KL1][KL1];
    double beta[kMaxFn], mu[kMaxFn];
    T ux[NTH], uy[NTH], uz[NTH];
    double rr[NTH], fc[NTH];
    int spec[NTH];
    alignas(16) double g[kAnlmJ][GW + 2];  // rows 16 B longer: a column's stores on distinct banks
};

// Radial indices per round of the feature epilogue: a of both species
// (FP64 pairs) and the round's feature segments must fit the Y_lm chunk
// buffer they reuse, kAnlmJ rows of n_lmh entries of anlm_buf_entry bytes.
// The compact tables (l_max <= 12) use 16-byte entries in both precisions
// (an FP32 chunk uses half: the mixed mode's epilogue then takes 7 radial
// indices per round at l_max 12, as FP64 does, instead of 3, and runs half
// the rounds and their barriers, with 4 workgroups per CU either way); the
// full tables keep entries of T2, since a larger buffer there would cost a
// workgroup per CU.
template <typename T, int KL1>
constexpr int anlm_buf_entry() {
    return KL1 == kSmallL1 ? 16 : static_cast<int>(sizeof(typename Vec2<T>::type));
}
__host__ __device__ inline int anlm_feature_round(int n_lmh, int w, int t2_bytes) {
    const int buf = kAnlmJ * n_lmh * t2_bytes;
    return buf / (2 * n_lmh * 16 + w * 8);
}

template <int NPT, int NTH, typename T, int KL1, int GW>
__global__ __launch_bounds__(NTH) void k_anlm(DevModel M, DevSystem S, int ngroup, int feat) {
    constexpr int kAnlmThreads = NTH;
    constexpr int kAnlmParts = NTH / kAnlmJ;
    using T2 = typename Vec2<T>::type;
    __shared__ AnlmLds<T, NTH, KL1, GW> L;
    extern __shared__ __attribute__((aligned(16))) unsigned char anlm_dyn[];
    T2* ys = reinterpret_cast<T2*>(anlm_dyn);  // [kAnlmJ][n_lmh]
    const int slot = xcd_slot(blockIdx.x, S.nblocks);
    if (slot < 0) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) return;
    const int tid = threadIdx.x;
    const int nfn = M.n_fn, nlmh = M.n_lmh, maxl = M.maxl;
    {
        const YTab<T>* yt = ytab_of<T>(M);
        for (int k = tid; k < KL1 * KL1; k += kAnlmThreads) {
            const int l = k / KL1, mm = k % KL1;
            L.alm[l][mm] = yt->alm[l][mm];
            L.ainv[l][mm] = yt->ainv[l][mm];
        }
        if (tid <= KL1) L.diag[tid] = yt->diag[tid];
        if (tid < kMaxFn) {
            L.beta[tid] = tid < nfn ? M.beta[tid] : 0.0;
            L.mu[tid] = tid < nfn ? M.mu[tid] : 0.0;
        }
    }
    const int p0 = S.poff[atom], p1 = S.pend[atom];
    // phase-2 role
    const int lm2 = tid % nlmh, grp = tid / nlmh;
    const bool act2 = grp < ngroup;
    const int nbase = (blockIdx.y * ngroup + grp) * NPT;
    double a0r[NPT], a0i[NPT], a1r[NPT], a1i[NPT];
#pragma unroll
    for (int k = 0; k < NPT; ++k) a0r[k] = a0i[k] = a1r[k] = a1i[k] = 0.0;
    // phase-1 role
    // lanes 16 apart share a column m: a 16-lane group takes one m for the
    // chunk's 16 neighbors, so a wave's lanes run at most four different
    // column lengths (the waves of the short columns finish early instead of
    // every wave waiting for m = 0), the recursion constants are broadcasts,
    // and a group's Y stores fall on distinct banks (rows n_lmh apart)
    const int j1 = tid % kAnlmJ, part = tid / kAnlmJ;
    const double pir = kPiRef / M.cutoff;
    __syncthreads();
    for (int b0 = p0; b0 < p1; b0 += kAnlmThreads) {
        const int nb = min(kAnlmThreads, p1 - b0);
        if (tid < nb) {
            const PairDel g = pair_del(M, S, atom, b0 + tid);
            int spec = -1;
            if (g.valid) {
                spec = S.elem[S.pnbr[b0 + tid]];
                const double ir = 1.0 / g.r;
                L.ux[tid] = static_cast<T>(g.dx * ir);
                L.uy[tid] = static_cast<T>(g.dy * ir);
                L.uz[tid] = static_cast<T>(g.dz * ir);
                L.rr[tid] = g.r;
                L.fc[tid] = 0.5 * (cos(pir * g.r) + 1.0);
            }
            L.spec[tid] = spec;
        }
        __syncthreads();
        for (int c0 = 0; c0 < nb; c0 += kAnlmJ) {
            const int q = c0 + j1;
            if (q < nb && L.spec[q] >= 0) {
                // radial functions in FP64 in both modes: FP32 cannot carry
                // the exponent beta (r - mu)^2 to the digits exp needs
                const double r = L.rr[q], fcq = L.fc[q];
                for (int n = part; n < GW; n += kAnlmParts) {
                    const double d = n < nfn ? r - L.mu[n] : 0.0;
                    L.g[j1][n] = n < nfn ? exp(-L.beta[n] * (d * d)) * fcq : 0.0;
                }
                const T x = L.ux[q], y = L.uy[q], z = L.uz[q];
                T2* yrow = ys + j1 * nlmh;
                // columns in boustrophedon order over the parts (part, then
                // 2 P - 1 - part, ...): the second round's short columns go to
                // the last parts, whose waves hold the short first-round ones
                for (int r = 0, m = part; m <= maxl; ++r, m = (r & 1) ? (r + 1) * kAnlmParts - 1 - part : r * kAnlmParts + part) {
                    // (-1)^m conj(u)^m by binary powering
                    T pr = (m & 1) ? T(-1) : T(1), pi = T(0);
                    T br = x, bi = -y;
                    for (int e = m; e; e >>= 1) {
                        if (e & 1) {
                            const T t = pr * br - pi * bi;
                            pi = pr * bi + pi * br;
                            pr = t;
                        }
                        const T t2 = br * br - bi * bi;
                        bi = T(2) * br * bi;
                        br = t2;
                    }
                    T q2 = T(0), q1 = L.diag[m];
                    int idx = m * (m + 1) / 2;  // lmh(m, -m)
                    yrow[idx] = T2{q1 * pr, q1 * pi};
                    // two steps per trip: half the loop's exec-mask and index
                    // work, and no register rotation of q1, q2
#pragma unroll 2
                    for (int l = m + 1; l <= maxl; ++l) {
                        const T qn = L.alm[l][m] * (z * q1 - L.ainv[l][m] * q2);
                        idx += l + 1;  // lmh(l, -m) = lmh(l-1, -m) + l + 1
                        yrow[idx] = T2{qn * pr, qn * pi};
                        q2 = q1;
                        q1 = qn;
                    }
                }
            }
            __syncthreads();
            const int nc = min(kAnlmJ, nb - c0);
            // species of the chunk's neighbors, one per lane: each neighbor's
            // branch reads a register (readlane), not an LDS round trip
            // (nc <= kAnlmJ <= 64, so lanes 0..nc-1 of every wave hold them)
            const int wl = tid & 63;
            const int specv = wl < nc ? L.spec[c0 + wl] : -1;
            if (act2) {
                for (int j = 0; j < nc; ++j) {
                    const int spec = __builtin_amdgcn_readlane(specv, j);
                    if (spec < 0) continue;
                    const T2 yv = ys[j * nlmh + lm2];
                    const double yr = static_cast<double>(yv.x), yi = static_cast<double>(yv.y);
                    // g_n by 16-byte LDS reads (NPT is even, nbase a multiple of NPT)
                    const double2* gj2 = reinterpret_cast<const double2*>(&L.g[j][nbase]);
                    double gj[NPT];
#pragma unroll
                    for (int h = 0; h < NPT / 2; ++h) {
                        const double2 v = gj2[h];
                        gj[2 * h] = v.x;
                        gj[2 * h + 1] = v.y;
                    }
                    if (spec == 0) {
#pragma unroll
                        for (int k = 0; k < NPT; ++k) {
                            a0r[k] = fma(gj[k], yr, a0r[k]);
                            a0i[k] = fma(gj[k], yi, a0i[k]);
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < NPT; ++k) {
                            a1r[k] = fma(gj[k], yr, a1r[k]);
                            a1i[k] = fma(gj[k], yi, a1i[k]);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    if (act2) {
        const int blk = slot / kBlock, lane = slot % kBlock;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int n = nbase + k;
            if (n < nfn) {
                S.A[a_index(M, blk, n, 0, lm2, 0, lane)] = a0r[k];
                S.A[a_index(M, blk, n, 0, lm2, 1, lane)] = a0i[k];
                if (M.n_type > 1) {
                    S.A[a_index(M, blk, n, 1, lm2, 0, lane)] = a1r[k];
                    S.A[a_index(M, blk, n, 1, lm2, 1, lane)] = a1i[k];
                }
            }
        }
    }
    if (!feat) return;
    // Order-1/2 polynomial features of this centre (the order-1/2 feature
    // pass's d, formed here from the a_nlm the workgroup holds): per round
    // of R radial indices, a of both species staged in LDS [n][s][lm] over
    // the Y_lm chunk buffer, thread per (feature, n), features gathered in
    // an LDS row segment [n][rank - lo] (zero where no order-1/2 family
    // writes) and stored as one contiguous piece of the slot's compact row.
    const int blk = slot / kBlock;
    const int t = S.block_type[blk];
    const int lo = M.prange[t][0], w = M.prange[t][1] - lo;
    const int f0 = M.fq12_off[t], nfe = M.fq12_off[t + 1] - f0;
    const int R = anlm_feature_round(nlmh, w, anlm_buf_entry<T, KL1>());
    const int nw0 = blockIdx.y * ngroup * NPT, nw1 = min(nw0 + ngroup * NPT, nfn);
    double2* as = reinterpret_cast<double2*>(anlm_dyn);
    for (int r0 = nw0; r0 < nw1; r0 += R) {
        const int rn = min(R, nw1 - r0);
        double* tile = reinterpret_cast<double*>(as + static_cast<size_t>(rn) * 2 * nlmh);
        __syncthreads();  // the chunk buffer (or the last round) is free
        if (act2) {
#pragma unroll
            for (int k = 0; k < NPT; ++k) {
                const int n = nbase + k;
                if (n >= r0 && n < r0 + rn) {
                    as[((n - r0) * 2 + 0) * nlmh + lm2] = double2{a0r[k], a0i[k]};
                    as[((n - r0) * 2 + 1) * nlmh + lm2] = double2{a1r[k], a1i[k]};
                }
            }
        }
        for (int e = tid; e < rn * w; e += kAnlmThreads) tile[e] = 0.0;
        __syncthreads();
        // item = (feature, n) with n fastest: the features are listed by L,
        // so a wave's items have nearly the same number of terms
        for (int it = tid; it < nfe * rn; it += kAnlmThreads) {
            const int fi = it / rn, nl = it - fi * rn;
            const FeatQ12 f = M.fq12[f0 + fi];
            const double* cp = M.coef + f.coef;
            const double2* x = as + (nl * 2 + f.s0) * nlmh;
            double d;
            if (f.order == 1) {
                d = cp[0] * x[0].x;
            } else {
                // the order-1/2 pass's pair-summed form, same order of terms
                const double2* y = as + (nl * 2 + f.s1) * nlmh;
                const int L = f.L, b = L * (L + 1) / 2 + L;  // lmh(L, 0); lmh(L, -k) = b - k
                d = 0.0;
                for (int k = 0; k <= L; ++k) {
                    const double sg = (k & 1) ? -1.0 : 1.0;
                    const double2 xv = x[b - k], yv = y[b - k];
                    const double cm = cp[L - k], cpk = cp[L + k];
                    const double re = sg * (xv.x * yv.x + xv.y * yv.y);
                    d = fma(k == 0 ? cm : cm + cpk, re, d);
                }
            }
            tile[nl * w + f.col] = d;
        }
        __syncthreads();
        // columns r0 w .. (r0 + rn) w of the slot's compact row
        double* drow = S.D + static_cast<size_t>(slot) * M.n_pr + static_cast<size_t>(r0) * w;
        for (int e = tid; e < rn * w; e += kAnlmThreads) drow[e] = tile[e];
    }
}

// ---------------------------------------------------------------- a_nlm on the matrix cores
// The same sum as a product per chunk of 16 neighbors:
//   a_s[n][(lm, re/im)] += sum_k At[k][n] Y[k][(lm, re/im)],
// k = the chunk's neighbors with species s, on v_mfma_f64_16x16x4f64 (A[i][k]
// from lane 16 k + i, B[k][j] from lane 16 k + j, result register r of lane
// l is D[l / 16 + 4 r][l % 16]).  Phase 1 puts each chunk's neighbors in LDS
// sorted by species (species 0 first), so a step of four neighbors feeds one
// species' accumulators (two for the step that straddles the boundary, with
// the other species' columns masked to zero) and no lane branches on species.
// One workgroup of four waves per centre; wave w owns the column tiles
// w, w + 4, ... and both species' accumulators for them (MT tiles of 16 radial
// indices each).  Per-pair geometry, g_n and Y_lm are computed as in k_anlm
// (radial functions in FP64, Y_lm in T); the product and its accumulation
// are FP64.
constexpr int kAmThreads = 256;
constexpr int kAmJ = 16;

template <typename T, int MT, int NTW>
struct AmLds {
    static constexpr int kCols = 64 * NTW;      // 4 waves x NTW tiles x 16 columns
    static constexpr int kYPitch = kCols + 2;   // doubles; offsets the rows' banks
    static constexpr int kAPitch = 16 * MT + 2;
    T diag[kSmallL1 + 1];
    T alm[kSmallL1][kSmallL1];
    T ainv[kSmallL1][kSmallL1];
    double beta[kMaxFn], mu[kMaxFn];
    T ux[kAmThreads], uy[kAmThreads], uz[kAmThreads];
    double rr[kAmThreads], fc[kAmThreads];
    int spec[kAmThreads];
    alignas(16) double At[kAmJ][kAPitch];   // [sorted neighbor][n]
    alignas(16) double Y[kAmJ][kYPitch];    // [sorted neighbor][(lm, re/im)]
};

template <typename T, int MT, int NTW>
__global__ __launch_bounds__(kAmThreads) void k_anlm_mfma(DevModel M, DevSystem S) {
    using L_t = AmLds<T, MT, NTW>;
    constexpr int kCols = L_t::kCols;
    __shared__ L_t L;
    const int slot = xcd_slot(blockIdx.x, S.nblocks);
    if (slot < 0) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nfn = M.n_fn, nlmh = M.n_lmh, maxl = M.maxl;
    {
        const YTab<T>* yt = ytab_of<T>(M);
        for (int k = tid; k < kSmallL1 * kSmallL1; k += kAmThreads) {
            const int l = k / kSmallL1, mm = k % kSmallL1;
            L.alm[l][mm] = yt->alm[l][mm];
            L.ainv[l][mm] = yt->ainv[l][mm];
        }
        if (tid <= kSmallL1) L.diag[tid] = yt->diag[tid];
        if (tid < kMaxFn) {
            L.beta[tid] = tid < nfn ? M.beta[tid] : 0.0;
            L.mu[tid] = tid < nfn ? M.mu[tid] : 0.0;
        }
        // Y columns past 2 n_lmh stay zero for the whole kernel
        for (int k = tid; k < kAmJ * kCols; k += kAmThreads) L.Y[k / kCols][k % kCols] = 0.0;
    }
    typedef double dbl4 __attribute__((ext_vector_type(4)));
    dbl4 acc[2][MT][NTW];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int t = 0; t < NTW; ++t) acc[s][i][t] = dbl4{0.0, 0.0, 0.0, 0.0};
    const int p0 = S.poff[atom], p1 = S.pend[atom];
    const double pir = kPiRef / M.cutoff;
    // phase 1: lanes 16 apart share a column m (part), the 16 lanes of a
    // group take the chunk's 16 neighbors (j1), as in k_anlm
    const int j1 = tid & 15, part = tid >> 4;
    __syncthreads();
    for (int b0 = p0; b0 < p1; b0 += kAmThreads) {
        const int nb = min(kAmThreads, p1 - b0);
        if (tid < nb) {
            const PairDel g = pair_del(M, S, atom, b0 + tid);
            int spec = -1;
            if (g.valid) {
                spec = S.elem[S.pnbr[b0 + tid]];
                const double ir = 1.0 / g.r;
                L.ux[tid] = static_cast<T>(g.dx * ir);
                L.uy[tid] = static_cast<T>(g.dy * ir);
                L.uz[tid] = static_cast<T>(g.dz * ir);
                L.rr[tid] = g.r;
                L.fc[tid] = 0.5 * (cos(pir * g.r) + 1.0);
            }
            L.spec[tid] = spec;
        }
        __syncthreads();
        for (int c0 = 0; c0 < nb; c0 += kAmJ) {
            const int nc = min(kAmJ, nb - c0);
            // sorted order of the chunk: species 0, then species 1; pairs
            // beyond the cutoff get no place
            const int specv = lane < nc ? L.spec[c0 + lane] : -1;
            const uint64_t m0 = __ballot(specv == 0), m1 = __ballot(specv == 1);
            const int n0 = __popcll(m0), nval = n0 + __popcll(m1);
            // phase 1: g_n and Y of neighbor j1 into its sorted row
            {
                const int q = c0 + j1;
                const int sp = j1 < nc ? L.spec[q] : -1;
                const uint64_t below = (1ull << j1) - 1ull;
                const int k = sp == 0 ? __popcll(m0 & below) : (sp == 1 ? n0 + __popcll(m1 & below) : -1);
                if (k >= 0) {
                    const double r = L.rr[q], fcq = L.fc[q];
                    for (int n = part; n < 16 * MT; n += 16) {
                        const double d = n < nfn ? r - L.mu[n] : 0.0;
                        L.At[k][n] = n < nfn ? exp(-L.beta[n] * (d * d)) * fcq : 0.0;
                    }
                    const T x = L.ux[q], y = L.uy[q], z = L.uz[q];
                    double* yrow = L.Y[k];
                    for (int m = part; m <= maxl; m += 16) {
                        T pr = (m & 1) ? T(-1) : T(1), pi = T(0);  // (-1)^m conj(u)^m
                        T br = x, bi = -y;
                        for (int e = m; e; e >>= 1) {
                            if (e & 1) {
                                const T t = pr * br - pi * bi;
                                pi = pr * bi + pi * br;
                                pr = t;
                            }
                            const T t2 = br * br - bi * bi;
                            bi = T(2) * br * bi;
                            br = t2;
                        }
                        T q2 = T(0), q1 = L.diag[m];
                        int idx = m * (m + 1) / 2;  // lmh(m, -m)
                        yrow[2 * idx] = static_cast<double>(q1 * pr);
                        yrow[2 * idx + 1] = static_cast<double>(q1 * pi);
#pragma unroll 2
                        for (int l = m + 1; l <= maxl; ++l) {
                            const T qn = L.alm[l][m] * (z * q1 - L.ainv[l][m] * q2);
                            idx += l + 1;  // lmh(l, -m) = lmh(l-1, -m) + l + 1
                            yrow[2 * idx] = static_cast<double>(qn * pr);
                            yrow[2 * idx + 1] = static_cast<double>(qn * pi);
                            q2 = q1;
                            q1 = qn;
                        }
                    }
                }
                // rows past the chunk's valid neighbors are zero
                for (int e = tid; e < (kAmJ - nval) * 16 * MT; e += kAmThreads)
                    L.At[nval + e / (16 * MT)][e % (16 * MT)] = 0.0;
                for (int e = tid; e < (kAmJ - nval) * 2 * nlmh; e += kAmThreads)
                    L.Y[nval + e / (2 * nlmh)][e % (2 * nlmh)] = 0.0;
            }
            __syncthreads();
            // phase 2: four steps of four sorted neighbors
            for (int ks = 0; ks < kAmJ / 4; ++ks) {
                const int kb = 4 * ks;
                if (kb >= nval) break;
                const int kr = kb + (lane >> 4);
                const bool s0 = kb < n0, s1 = kb + 3 >= n0;  // species present in the step
                double a0[MT], a1[MT];
#pragma unroll
                for (int i = 0; i < MT; ++i) {
                    const double a = L.At[kr][16 * i + (lane & 15)];
                    a0[i] = kr < n0 ? a : 0.0;
                    a1[i] = kr < n0 ? 0.0 : a;
                }
#pragma unroll
                for (int t = 0; t < NTW; ++t) {
                    const double b = L.Y[kr][(wave * NTW + t) * 16 + (lane & 15)];
#pragma unroll
                    for (int i = 0; i < MT; ++i) {
                        if (s0) acc[0][i][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[i], b, acc[0][i][t], 0, 0, 0);
                        if (s1) acc[1][i][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[i], b, acc[1][i][t], 0, 0, 0);
                    }
                }
            }
            __syncthreads();
        }
    }
    const int blk = slot / kBlock, sl = slot % kBlock;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (s >= M.n_type) break;
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int t = 0; t < NTW; ++t) {
                const int col = (wave * NTW + t) * 16 + (lane & 15);
                if (col >= 2 * nlmh) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int n = 16 * i + (lane >> 4) + 4 * r;
                    if (n < nfn) S.A[a_index(M, blk, n, s, col >> 1, col & 1, sl)] = acc[s][i][t][r];
                }
            }
    }
}

// ---------------------------------------------------------------- forces
// One workgroup per centre; G of the centre (both species) staged in LDS as
// T [s][lm][n][re,im] with n padded to NF (zeros).  Thread = neighbor pair;
// Y_lm and grad Y_lm by the recursion above (coefficients from LDS, read
// as broadcasts), contracted with
//   Gp_lm = sum_n g'_n G_nlm,  Gq_lm = sum_n g_n G_nlm
//   dE/d(del) = sum_lm w_m Re(Gp Y) rhat + w_m Re(Gq . grad Y),  w_0 = 1, w_{m<0} = 2
// grad Y_l^m * r = {(-B u^{m+1} + E u^{m-1})/2, (i/2)(B u^{m+1} + E u^{m-1}), -A u^m},
//   A = m z Q + (1-z^2) c Q', B = m Q - z c Q', E = m (1+z^2) Q + z (1-z^2) c Q',
//   Q = Qb_l^m, Q' = Qb_l^{m+1}, c = sqrt((l+m+1)(l-m)); for m = 0 the
//   singular terms cancel to c Q' (z x, z y, -(1-z^2)).
template <typename T, int KL1>
struct ForceLds {
    double beta[kMaxFn], mu[kMaxFn];
    T diag[KL1 + 1];
    T alm[KL1][KL1];
    T ainv[KL1][KL1];
    T grad[KL1][KL1];
    double trig[kTrigN + 1];  // sin / cos series coefficients, then pi / r_c
};

// Register budget for occupancy: FP64 at most 168 VGPRs (3 waves per SIMD;
// unconstrained it takes 214 and runs 2 waves), FP32 at most 128 (4 waves).
// The spills this costs are cheaper than the lost latency hiding: 28.0 ->
// 19.6 ms (FP64) and 14.2 -> 12.9 ms (FP32) at 256k atoms with 192 threads,
// where the workgroups' G tiles still fit the LDS.  Models whose G tile does
// not fit (l_max and n_fn near their limits) read G from HBM instead (GL =
// false; every thread of the workgroup reads the same entries, served by the
// caches), and models with more than 16 radial functions keep the register
// arrays they need (no occupancy cap).
template <int NF, int NT, typename T, bool GL, int KL1>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NF > 16 ? 1 : (sizeof(T) == 4 ? 4 : 3)))) void k_force(DevModel M, DevSystem S) {
    extern __shared__ __attribute__((aligned(16))) unsigned char force_dyn[];
    T* Gs = reinterpret_cast<T*>(force_dyn);  // [ns][n_lmh][NF][2]
    __shared__ ForceLds<T, KL1> C;
    __shared__ double vred[6][NT / 64];
    const int slot = xcd_slot(blockIdx.x, S.nblocks);
    if (slot < 0) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) return;
    const int nfn = M.n_fn, ns = M.n_type, nlmh = M.n_lmh, maxl = M.maxl;
    const int blk = slot / kBlock, lane = slot % kBlock;
    // four tile entries per step, read without a branch (padding columns
    // n >= n_fn read an in-range entry and store zero), so that their HBM /
    // L2 reads are in flight together
#pragma unroll 4
    for (int k = threadIdx.x; GL && k < ns * nlmh * NF; k += NT) {
        const int n = k % NF;
        const int lm = (k / NF) % nlmh;
        const int s = k / (NF * nlmh);
        const int nr = n < nfn ? n : nfn - 1;
        const double gr = S.G[a_index(M, blk, nr, s, lm, 0, lane)];
        const double gi = S.G[a_index(M, blk, nr, s, lm, 1, lane)];
        Gs[2 * k] = n < nfn ? static_cast<T>(gr) : T(0);
        Gs[2 * k + 1] = n < nfn ? static_cast<T>(gi) : T(0);
    }
    {
        const YTab<T>* yt = ytab_of<T>(M);
        for (int k = threadIdx.x; k < KL1 * KL1; k += NT) {
            const int l = k / KL1, mm = k % KL1;
            C.alm[l][mm] = yt->alm[l][mm];
            C.ainv[l][mm] = yt->ainv[l][mm];
            C.grad[l][mm] = yt->grad[l][mm];
        }
        if (threadIdx.x <= KL1) C.diag[threadIdx.x] = yt->diag[threadIdx.x];
        if (threadIdx.x < kTrigN) C.trig[threadIdx.x] = trig_coef(threadIdx.x);
        if (threadIdx.x == kTrigN) C.trig[kTrigN] = kPiRef / M.cutoff;
        if (threadIdx.x < kMaxFn) {
            C.beta[threadIdx.x] = threadIdx.x < nfn ? M.beta[threadIdx.x] : 0.0;
            C.mu[threadIdx.x] = threadIdx.x < nfn ? M.mu[threadIdx.x] : 0.0;
        }
    }
    __syncthreads();
    const int p0 = S.poff[atom], p1 = S.pend[atom];
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += NT) {
        const PairDel pg = pair_del(M, S, atom, p);
        double fx = 0.0, fy = 0.0, fz = 0.0;
        double ex = pg.dx, ey = pg.dy, ez = pg.dz;  // del for the virial
        if (pg.valid) {
            const int spec = S.elem[S.pnbr[p]];
            const T* Gsp = Gs + (size_t)spec * nlmh * NF * 2;
            const double ird = 1.0 / pg.r;
            const T x = static_cast<T>(pg.dx * ird), y = static_cast<T>(pg.dy * ird),
                    z = static_cast<T>(pg.dz * ird);
            // radial functions and the cutoff function in FP64 in both modes; pi / r_c
            // and the series coefficients are read from LDS here rather than held
            lds_cdouble* tq = lds_opaque(C.trig);
            const double pir = tq[kTrigN];
            const double arg = pir * pg.r;
            double sa, ca;
            sin_cos(arg, tq, &sa, &ca);
            const double fc = 0.5 * (ca + 1.0);
            const double fcd = -0.5 * pir * sa;
            T gn[NF], gd[NF];
#pragma unroll
            for (int n = 0; n < NF; ++n) {
                // padded n >= n_fn meet zero G columns
                const double dd = pg.r - C.mu[n];
                const double be = C.beta[n];
                const double b = exp(-be * (dd * dd));
                gn[n] = static_cast<T>(b * fc);
                gd[n] = static_cast<T>(b * (fcd - 2.0 * be * dd * fc));
                // one radial function at a time: interleaving the twelve
                // exponentials costs more registers than it hides latency
                __builtin_amdgcn_sched_barrier(0);
            }
            // The gradient terms are linear in q = Qb_l^m and c Q' with
            // coefficients that depend on m only, so each m column keeps three
            // complex sums over l (in T)
            //   Sp = sum_l q Gp,  Sq = sum_l q Gq,  Sd = sum_l c Q' Gq
            // and folds them into er, t (FP64 in both modes) once per column:
            //   er += w_m (-1)^m Re(Sp conj(u)^m)
            //   m = 0: t += Re(Sd) (z x, z y, -(1-z^2))
            //   m > 0: SA = m z Sq + (1-z^2) Sd, SB = m Sq - z Sd,
            //          SE = m (1+z^2) Sq + z (1-z^2) Sd,
            //          t += w_m (-1)^m (Re(-SB u+ + SE u-) / 2, Im(SB u+ + SE u-) / 2, -Re(SA u0))
            //   with u0, u+, u- = conj(u)^m, conj(u)^(m+1), conj(u)^(m-1).
            double er = 0.0;                        // sum w Re(Gp Y)
            double tx = 0.0, ty = 0.0, tz = 0.0;    // sum w Re(Gq grad Y) * r
            const double xd = static_cast<double>(x), yd = static_cast<double>(y), zd = static_cast<double>(z);
            const double zzd = 1.0 - zd * zd;
            T cr = T(1), ci = T(0);  // conj(u)^m
            T mr = T(0), mi = T(0);  // conj(u)^(m-1)
#pragma unroll 1
            for (int m = 0; m <= maxl; ++m) {
                const T pr = cr * x + ci * y, pi = ci * x - cr * y;  // conj(u)^(m+1)
                T a1 = T(0), a2 = T(0);  // Qb_{l-1}^m, Qb_{l-2}^m
                T b1 = T(0), b2 = T(0);  // Qb_{l-1}^{m+1}, Qb_{l-2}^{m+1}
                T spr = T(0), spi = T(0), sqr = T(0), sqi = T(0), sdr = T(0), sdi = T(0);
                int lm = m * (m + 1) / 2;  // lmh(m, -m)
#pragma unroll 1
                for (int l = m; l <= maxl; ++l) {
                    // contraction over n first: independent of the recursion
                    T gpr = T(0), gpi = T(0), gqr = T(0), gqi = T(0);
                    // 16-byte LDS reads (one n in FP64, two in FP32)
                    const T* gl = Gsp + (size_t)lm * NF * 2;
                    T grv[2 * NF];
                    if constexpr (!GL) {
#pragma unroll
                        for (int n = 0; n < NF; ++n) {
                            grv[2 * n] = n < nfn ? static_cast<T>(S.G[a_index(M, blk, n, spec, lm, 0, lane)]) : T(0);
                            grv[2 * n + 1] = n < nfn ? static_cast<T>(S.G[a_index(M, blk, n, spec, lm, 1, lane)]) : T(0);
                        }
                    } else if constexpr (sizeof(T) == 8) {
                        const double2* g2 = reinterpret_cast<const double2*>(gl);
#pragma unroll
                        for (int n = 0; n < NF; ++n) {
                            const double2 v = g2[n];
                            grv[2 * n] = v.x;
                            grv[2 * n + 1] = v.y;
                        }
                    } else {
                        const float4* g4 = reinterpret_cast<const float4*>(gl);
#pragma unroll
                        for (int h = 0; h < NF / 2; ++h) {
                            const float4 v = g4[h];
                            grv[4 * h] = v.x;
                            grv[4 * h + 1] = v.y;
                            grv[4 * h + 2] = v.z;
                            grv[4 * h + 3] = v.w;
                        }
                    }
#pragma unroll
                    for (int n = 0; n < NF; ++n) {
                        const T grr = grv[2 * n], gii = grv[2 * n + 1];
                        gpr = fma(gd[n], grr, gpr);
                        gpi = fma(gd[n], gii, gpi);
                        gqr = fma(gn[n], grr, gqr);
                        gqi = fma(gn[n], gii, gqi);
                    }
                    const T q = (l == m) ? C.diag[m] : C.alm[l][m] * (z * a1 - C.ainv[l][m] * a2);
                    a2 = a1;
                    a1 = q;
                    T qp = T(0);  // Qb_l^{m+1}
                    if (l == m + 1)
                        qp = C.diag[m + 1];
                    else if (l > m + 1)
                        qp = C.alm[l][m + 1] * (z * b1 - C.ainv[l][m + 1] * b2);
                    b2 = b1;
                    b1 = qp;
                    const T cq = C.grad[l][m] * qp;
                    spr = fma(q, gpr, spr);
                    spi = fma(q, gpi, spi);
                    sqr = fma(q, gqr, sqr);
                    sqi = fma(q, gqi, sqi);
                    sdr = fma(cq, gqr, sdr);
                    sdi = fma(cq, gqi, sdi);
                    lm += l + 2;  // lmh(l+1, -m) = lmh(l, -m) + l + 2
                }
                const double ws = (m == 0) ? 1.0 : ((m & 1) ? -2.0 : 2.0);  // w_m (-1)^m
                const double c0r = static_cast<double>(cr), c0i = static_cast<double>(ci);
                const double Spr = static_cast<double>(spr), Spi = static_cast<double>(spi);
                const double Sdr = static_cast<double>(sdr), Sdi = static_cast<double>(sdi);
                er += ws * (Spr * c0r - Spi * c0i);
                if (m == 0) {
                    tx += Sdr * zd * xd;
                    ty += Sdr * zd * yd;
                    tz -= Sdr * zzd;
                } else {
                    const double Sqr = static_cast<double>(sqr), Sqi = static_cast<double>(sqi);
                    const double fm = static_cast<double>(m);
                    const double kaq = fm * zd, kbq = fm, keq = fm * (1.0 + zd * zd);
                    const double ked = zd * zzd;
                    const double sar = kaq * Sqr + zzd * Sdr, sai = kaq * Sqi + zzd * Sdi;
                    const double sbr = kbq * Sqr - zd * Sdr, sbi = kbq * Sqi - zd * Sdi;
                    const double ser = keq * Sqr + ked * Sdr, sei = keq * Sqi + ked * Sdi;
                    const double upr = static_cast<double>(pr), upi = static_cast<double>(pi);
                    const double umr = static_cast<double>(mr), umi = static_cast<double>(mi);
                    // SB u+ and SE u- (complex products)
                    const double bpr = sbr * upr - sbi * upi, bpi = sbr * upi + sbi * upr;
                    const double emr = ser * umr - sei * umi, emi = ser * umi + sei * umr;
                    tx += 0.5 * ws * (emr - bpr);
                    ty += 0.5 * ws * (bpi + emi);
                    tz -= ws * (sar * c0r - sai * c0i);
                }
                mr = cr;
                mi = ci;
                cr = pr;
                ci = pi;
            }
            // dE/d(del) = er * rhat + t / r.  In FP64, rhat is (x, y, z) itself and del
            // is r rhat within an ulp, so of the pair's geometry only r stays live across
            // the column loop (del and 1/r held there were spilled to scratch); the mixed
            // mode, whose unit vector is FP32, keeps del.
            if constexpr (sizeof(T) == 8) {
                double rl = pg.r;
                asm volatile("" : "+v"(rl));  // 1/r recomputed here, not kept from above
                const double ir = 1.0 / rl;
                fx = er * xd + tx * ir;
                fy = er * yd + ty * ir;
                fz = er * zd + tz * ir;
                ex = xd * rl;
                ey = yd * rl;
                ez = zd * rl;
            } else {
                const double ir = ird;
                fx = er * (pg.dx * ir) + tx * ir;
                fy = er * (pg.dy * ir) + ty * ir;
                fz = er * (pg.dz * ir) + tz * ir;
            }
        }
        S.gpair[3 * p] = fx;
        S.gpair[3 * p + 1] = fy;
        S.gpair[3 * p + 2] = fz;
        // virial: del (x) F_centre with F_centre = -dE/d(del)
        v0 -= ex * fx;
        v1 -= ey * fy;
        v2 -= ez * fz;
        v3 -= ex * fy;
        v4 -= ex * fz;
        v5 -= ey * fz;
    }
    // fixed-order reduction: butterfly within each wave, then over waves
    for (int off = 32; off > 0; off >>= 1) {
        v0 += __shfl_xor(v0, off);
        v1 += __shfl_xor(v1, off);
        v2 += __shfl_xor(v2, off);
        v3 += __shfl_xor(v3, off);
        v4 += __shfl_xor(v4, off);
        v5 += __shfl_xor(v5, off);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        vred[0][wv] = v0;
        vred[1][wv] = v1;
        vred[2][wv] = v2;
        vred[3][wv] = v3;
        vred[4][wv] = v4;
        vred[5][wv] = v5;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = 0.0;
        for (int k = 0; k < NT / 64; ++k) s += vred[threadIdx.x][k];
        S.vown[6 * atom + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------- per-pair energy share
// Thread per directed pair of the slot's centre (64 threads, FP64):
//   e_p = sum_{n,l,m<=0} w_m Re(G_{s,n,l,m} g_n(r) Y_lm(del)),  w_0 = 1, w_{m<0} = 2
// with G read from HBM (every lane of the workgroup reads the same entry),
// plus the pair's virial del (x) (-g_p) from the force pass.  With G holding
// the split adjoint, e_p + e_p' over the two directions of a pair is the
// reference's pair energy evdwl (pair_mlip_gtinv.cpp:240-253).
__global__ __launch_bounds__(64) void k_pair_energy(DevModel M, DevSystem S, double* __restrict__ epv) {
    const int slot = xcd_slot(blockIdx.x, S.nblocks);
    if (slot < 0) return;
    const int atom = S.atom_of_slot[slot];
    if (atom < 0) return;
    const int blk = slot / kBlock, lane = slot % kBlock;
    const int nfn = M.n_fn, maxl = M.maxl;
    const YTab<double>* yt = M.yt64;
    const double pir = kPiRef / M.cutoff;
    const int p0 = S.poff[atom], p1 = S.pend[atom];
    for (int64_t p = p0 + threadIdx.x; p < p1; p += 64) {
        const PairDel pg = pair_del(M, S, atom, p);
        double e = 0.0;
        double v[6] = {0, 0, 0, 0, 0, 0};
        if (pg.valid) {
            const int spec = S.elem[S.pnbr[p]];
            const double ir = 1.0 / pg.r;
            const double x = pg.dx * ir, y = pg.dy * ir, z = pg.dz * ir;
            const double fc = 0.5 * (cos(pir * pg.r) + 1.0);
            double gn[kMaxFn];
            for (int n = 0; n < kMaxFn; ++n) {
                const double dd = pg.r - M.mu[n];
                gn[n] = n < nfn ? exp(-M.beta[n] * (dd * dd)) * fc : 0.0;
            }
            double cr = 1.0, ci = 0.0;  // conj(u)^m
            for (int m = 0; m <= maxl; ++m) {
                double q1 = 0.0, q2 = 0.0, sr = 0.0, si = 0.0;
                int lm = m * (m + 1) / 2;  // lmh(m, -m)
                for (int l = m; l <= maxl; ++l) {
                    const double q = (l == m) ? yt->diag[m] : yt->alm[l][m] * (z * q1 - yt->ainv[l][m] * q2);
                    q2 = q1;
                    q1 = q;
                    double gr = 0.0, gi = 0.0;
                    for (int n = 0; n < nfn; ++n) {
                        gr = fma(gn[n], S.G[a_index(M, blk, n, spec, lm, 0, lane)], gr);
                        gi = fma(gn[n], S.G[a_index(M, blk, n, spec, lm, 1, lane)], gi);
                    }
                    sr = fma(q, gr, sr);
                    si = fma(q, gi, si);
                    lm += l + 2;  // lmh(l+1, -m)
                }
                const double ws = (m == 0) ? 1.0 : ((m & 1) ? -2.0 : 2.0);  // w_m (-1)^m
                e += ws * (sr * cr - si * ci);
                const double nr = cr * x + ci * y, ni = ci * x - cr * y;
                cr = nr;
                ci = ni;
            }
            const double gx = S.gpair[3 * p], gy = S.gpair[3 * p + 1], gz = S.gpair[3 * p + 2];
            v[0] = -pg.dx * gx;
            v[1] = -pg.dy * gy;
            v[2] = -pg.dz * gz;
            v[3] = -pg.dx * gy;
            v[4] = -pg.dx * gz;
            v[5] = -pg.dy * gz;
        }
        epv[7 * p] = e;
#pragma unroll
        for (int c = 0; c < 6; ++c) epv[7 * p + 1 + c] = v[c];
    }
}

int slot_grid(const DevSystem& s) { return 8 * kBlock * ((s.nblocks + 7) / 8); }

// radial indices per phase-2 thread of the a_nlm kernel, lm groups per
// pass and passes over the neighbors (grid.y)
int anlm_npt(const DevModel& m, int nth, int* ngroup, int* passes) {
    const int gmax = nth / m.n_lmh;
    for (int npt : {2, 4, 6, 8, 16}) {
        const int g = (m.n_fn + npt - 1) / npt;
        if (g <= gmax) {
            *ngroup = g;
            *passes = 1;
            return npt;
        }
    }
    *ngroup = gmax;
    *passes = (m.n_fn + 16 * gmax - 1) / (16 * gmax);
    return 16;
}

template <int NTH, typename T, int KL1, int GW>
hipError_t launch_anlm_nth(const DevModel& m, const DevSystem& s, int npt, int ngroup, int passes, int feat,
                           hipStream_t st) {
    static_assert(sizeof(typename Vec2<T>::type) <= anlm_buf_entry<T, KL1>(), "a Y_lm chunk fits the buffer");
    const size_t lds = static_cast<size_t>(anlm_buf_entry<T, KL1>()) * kAnlmJ * m.n_lmh;
    if (lds + sizeof(AnlmLds<T, NTH, KL1, GW>) > lds_limit()) return hipErrorInvalidConfiguration;
    const dim3 grid(slot_grid(s), passes), block(NTH);
    switch (npt) {
        case 2: hipLaunchKernelGGL((k_anlm<2, NTH, T, KL1, GW>), grid, block, lds, st, m, s, ngroup, feat); break;
        case 4: hipLaunchKernelGGL((k_anlm<4, NTH, T, KL1, GW>), grid, block, lds, st, m, s, ngroup, feat); break;
        case 6: hipLaunchKernelGGL((k_anlm<6, NTH, T, KL1, GW>), grid, block, lds, st, m, s, ngroup, feat); break;
        case 8: hipLaunchKernelGGL((k_anlm<8, NTH, T, KL1, GW>), grid, block, lds, st, m, s, ngroup, feat); break;
        default: hipLaunchKernelGGL((k_anlm<16, NTH, T, KL1, GW>), grid, block, lds, st, m, s, ngroup, feat); break;
    }
    return launched();
}

template <typename T, int MT, int NTW>
hipError_t launch_anlm_mfma_k(const DevModel& m, const DevSystem& s, hipStream_t st) {
    if (sizeof(AmLds<T, MT, NTW>) > lds_limit()) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_anlm_mfma<T, MT, NTW>), dim3(slot_grid(s)), dim3(kAmThreads), 0, st, m, s);
    return launched();
}

// the matrix-core a_nlm kernel takes l_max <= 12, two species and up to 32
// radial functions (MT = 1 or 2 tiles of 16), with at most three column
// tiles per wave (2 n_lmh <= 192 columns)
bool anlm_mfma_fits(const DevModel& m) {
    const int nt = (2 * m.n_lmh + 15) / 16;
    return m.maxl < kSmallL1 && m.n_type <= 2 && m.n_fn <= 32 && (nt + 3) / 4 <= 3;
}

template <typename T>
hipError_t launch_anlm_mfma_t(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int mt = (m.n_fn + 15) / 16, ntw = ((2 * m.n_lmh + 15) / 16 + 3) / 4;
    if (mt == 1) {
        if (ntw == 1) return launch_anlm_mfma_k<T, 1, 1>(m, s, st);
        if (ntw == 2) return launch_anlm_mfma_k<T, 1, 2>(m, s, st);
        return launch_anlm_mfma_k<T, 1, 3>(m, s, st);
    }
    if (ntw == 1) return launch_anlm_mfma_k<T, 2, 1>(m, s, st);
    if (ntw == 2) return launch_anlm_mfma_k<T, 2, 2>(m, s, st);
    return launch_anlm_mfma_k<T, 2, 3>(m, s, st);
}

template <typename T>
hipError_t launch_anlm_t(const DevModel& m, const DevSystem& s, int feat, hipStream_t st) {
    // phase 2 needs a thread per lm column: 192 threads up to l_max 18
    const int nth = m.n_lmh <= 192 ? 192 : 256;
    int ngroup = 1, passes = 1;
    const int npt = anlm_npt(m, nth, &ngroup, &passes);
    if (m.maxl < kSmallL1 && passes * ngroup * npt <= kSmallGW)
        return launch_anlm_nth<192, T, kSmallL1, kSmallGW>(m, s, npt, ngroup, passes, feat, st);
    if (nth == 192) return launch_anlm_nth<192, T, kL1, kLargeGW>(m, s, npt, ngroup, passes, feat, st);
    return launch_anlm_nth<256, T, kL1, kLargeGW>(m, s, npt, ngroup, passes, feat, st);
}

template <int NF, int NT, typename T, bool GL, int KL1>
hipError_t launch_force_k(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const size_t lds = GL ? sizeof(T) * 2 * m.n_type * m.n_lmh * NF : 0;
    if (lds + sizeof(ForceLds<T, KL1>) + sizeof(double) * 6 * (NT / 64) > lds_limit())
        return hipErrorInvalidConfiguration;
    if (GL) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_force<NF, NT, T, GL, KL1>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_force<NF, NT, T, GL, KL1>), dim3(slot_grid(s)), dim3(NT), lds, st, m, s);
    return launched();
}

// G tile staged in LDS when it fits beside the kernel's static LDS
template <typename T, int KL1>
bool force_tile_fits(const DevModel& m, int nf, int nt) {
    const size_t lds = sizeof(T) * 2 * m.n_type * m.n_lmh * nf;
    return lds + sizeof(ForceLds<T, KL1>) + sizeof(double) * 6 * (nt / 64) <= lds_limit();
}

template <int NT, typename T>
hipError_t launch_force_nt(const DevModel& m, const DevSystem& s, hipStream_t st) {
    const int nf = m.n_fn <= 4 ? 4 : m.n_fn <= 8 ? 8 : m.n_fn <= 12 ? 12 : m.n_fn <= 16 ? 16 : m.n_fn <= 24 ? 24 : 32;
    if (m.maxl < kSmallL1 && nf <= 16 && force_tile_fits<T, kSmallL1>(m, nf, NT)) {
        if (nf == 4) return launch_force_k<4, NT, T, true, kSmallL1>(m, s, st);
        if (nf == 8) return launch_force_k<8, NT, T, true, kSmallL1>(m, s, st);
        if (nf == 12) return launch_force_k<12, NT, T, true, kSmallL1>(m, s, st);
        return launch_force_k<16, NT, T, true, kSmallL1>(m, s, st);
    }
    // large models (l_max above 12, more than 16 radial functions or a G
    // tile beyond the LDS): one launch shape, full-size tables
    if (force_tile_fits<T, kL1>(m, nf, 128)) {
        switch (nf) {
            case 4: return launch_force_k<4, 128, T, true, kL1>(m, s, st);
            case 8: return launch_force_k<8, 128, T, true, kL1>(m, s, st);
            case 12: return launch_force_k<12, 128, T, true, kL1>(m, s, st);
            case 16: return launch_force_k<16, 128, T, true, kL1>(m, s, st);
            case 24: return launch_force_k<24, 128, T, true, kL1>(m, s, st);
            default: return launch_force_k<32, 128, T, true, kL1>(m, s, st);
        }
    }
    switch (nf) {
        case 4: return launch_force_k<4, 128, T, false, kL1>(m, s, st);
        case 8: return launch_force_k<8, 128, T, false, kL1>(m, s, st);
        case 12: return launch_force_k<12, 128, T, false, kL1>(m, s, st);
        case 16: return launch_force_k<16, 128, T, false, kL1>(m, s, st);
        case 24: return launch_force_k<24, 128, T, false, kL1>(m, s, st);
        default: return launch_force_k<32, 128, T, false, kL1>(m, s, st);
    }
}

template <typename T>
hipError_t launch_force_t(const DevModel& m, const DevSystem& s, int threads, hipStream_t st) {
    switch (threads) {
        case 64: return launch_force_nt<64, T>(m, s, st);
        case 192: return launch_force_nt<192, T>(m, s, st);
        case 256: return launch_force_nt<256, T>(m, s, st);
        default: return launch_force_nt<128, T>(m, s, st);
    }
}

}  // namespace

void build_ylm_tables(YTab<double>& t64, YTab<float>& t32) {
    long double dg = 1.0L / sqrtl(4.0L * 3.14159265358979323846264338327950288L);
    for (int m = 0; m < kL1 + 1; ++m) {
        if (m > 0) dg *= -sqrtl((2.0L * m + 1.0L) / (2.0L * m));
        t64.diag[m] = static_cast<double>(dg);
    }
    for (int l = 0; l < kL1; ++l)
        for (int m = 0; m < kL1; ++m) {
            long double a = 0.0L, ai = 0.0L, g = 0.0L;
            if (l == m + 1) {
                a = sqrtl(2.0L * m + 3.0L);
            } else if (l >= m + 2) {
                a = sqrtl((4.0L * l * l - 1.0L) / ((long double)l * l - (long double)m * m));
                const long double a1 =
                    sqrtl((4.0L * (l - 1) * (l - 1) - 1.0L) / ((long double)(l - 1) * (l - 1) - (long double)m * m));
                ai = 1.0L / a1;
            }
            if (m <= l) g = sqrtl((long double)(l + m + 1) * (l - m));
            t64.alm[l][m] = static_cast<double>(a);
            t64.ainv[l][m] = static_cast<double>(ai);
            t64.grad[l][m] = static_cast<double>(g);
        }
    for (int m = 0; m < kL1 + 1; ++m) t32.diag[m] = static_cast<float>(t64.diag[m]);
    for (int l = 0; l < kL1; ++l)
        for (int m = 0; m < kL1; ++m) {
            t32.alm[l][m] = static_cast<float>(t64.alm[l][m]);
            t32.ainv[l][m] = static_cast<float>(t64.ainv[l][m]);
            t32.grad[l][m] = static_cast<float>(t64.grad[l][m]);
        }
}

hipError_t launch_anlm(const DevModel& m, const DevSystem& s, int precision, hipStream_t st, int kind, bool feat) {
    if (s.nblocks == 0) return hipSuccess;
    if (feat && !anlm_forms_features(m, precision, kind)) return hipErrorInvalidValue;
    if (kind == kAnlmMfma && anlm_mfma_fits(m))
        return precision == kMixed ? launch_anlm_mfma_t<float>(m, s, st) : launch_anlm_mfma_t<double>(m, s, st);
    const int f = feat ? 1 : 0;
    return precision == kMixed ? launch_anlm_t<float>(m, s, f, st) : launch_anlm_t<double>(m, s, f, st);
}

bool anlm_has_mfma(const DevModel& m) { return anlm_mfma_fits(m); }

// The register kernel forms the order-1/2 features when a round of at least
// one radial index fits its chunk buffer for every species' row segment.
bool anlm_forms_features(const DevModel& m, int precision, int kind) {
    if (kind == kAnlmMfma && anlm_mfma_fits(m)) return false;
    if (m.pair_model || m.n_p <= 0) return false;
    // the smallest buffer either table size may use (T2 entries): a round
    // that fits it fits the compact tables' larger one too
    const int t2 = precision == kMixed ? 8 : 16;
    for (int t = 0; t < m.n_type && t < 2; ++t)
        if (anlm_feature_round(m.n_lmh, m.prange[t][1] - m.prange[t][0], t2) < 1) return false;
    return true;
}

hipError_t launch_pair_energy(const DevModel& m, const DevSystem& s, double* epv, hipStream_t st) {
    if (s.nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_energy, dim3(slot_grid(s)), dim3(64), 0, st, m, s, epv);
    return launched();
}

hipError_t launch_force(const DevModel& m, const DevSystem& s, int precision, int threads, hipStream_t st) {
    if (s.nblocks == 0) return hipSuccess;
    return precision == kMixed ? launch_force_t<float>(m, s, threads, st) : launch_force_t<double>(m, s, threads, st);
}

}  // namespace mlip
