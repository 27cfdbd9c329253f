// Training outputs of the gtinv model on the device: the design matrices a
// fit needs (MLIPModel::prepare_features, src/USER-MLIP/mlipkk.cpp:461-470,
// getters :1322-1381).  For every atom the structural features
// X[i][f] = d_f^(i) (all n_des, f = n * n_gtinv + itc) and the polynomial
// features P[i][p] = prod_{f in p} X[i][f]; the force features
// FF[k][p] = -dX_p/dr_k with X_p = sum_i P[i][p]; per structure the stress
// features SF[p] = sum over pairs del (x) (force term) in the order
// xx, yy, zz, yz, zx, xy (mlipkk.cpp:1260-1271).  With the regression
// weights w: sum_p w_p X_p = E, sum_p w_p FF[k][p] = F_k and sum_p w_p SF[p]
// is the virial.
//
// Kernels (a_nlm from k_anlm in FP64 first):
//   k_tr_feat    thread per (slot lane, feature): CG terms over the lane's
//                a_nlm rows (lane-contiguous, coalesced)
//   k_tr_poly    thread per (atom, coefficient)
//   k_tr_dirj    workgroup per centre: its a_nlm in LDS; per chunk of its
//                pairs the geometry, g_n, g'_n and Y_lm, grad Y_lm in LDS;
//                then thread per (pair, itc, n): dirj = d d_f^(i)/d del_q
//                (mlipkk.cpp:992-1157 restated for a full list: each
//                directed pair carries its own centre's derivative)
//   k_tr_force   thread per (atom, coefficient): the atom's gather list in
//                pair order, g_qp = sum_nu prod_{nu' != nu} X[cen][f_nu'] dirj[q][f_nu],
//                -g on the centre, +g on the neighbor (fixed order)
//   k_tr_stress  workgroup per (structure, coefficient): del (x) (-g) over
//                the structure's pairs, fixed-order tree reduction
// All FP64.  Work and memory scale with pairs x n_des (dirj) and atoms x
// n_coeff (force features): a training-sized structure or batch, not the
// 256k-atom evaluation workload.
#include <climits>
#include <cstring>

#include "device_common.h"
#include "engine_internal.h"

namespace mlip {

struct TrItc {
    int32_t tbeg, tend;   // term range of the itc's irrep
    int32_t order;
    int32_t spec[2][4];   // [centre type][factor] -> neighbor species, -1: absent
};

struct TrTerm {
    double c;
    // per factor: lmh index of (l, -|m|); bit 16: m > 0 (conjugate);
    // bit 17: m > 0 and odd (negate)
    int32_t code[4];
};

struct TrainWork {
    bool built = false;
    DBuf<TrItc> itc;
    DBuf<TrTerm> term;
    DBuf<int32_t> rows;   // n_coeff x 3 feature ids (-1 padded)
    DBuf<double> X, P, FF, SF, dirj;
    DBuf<int32_t> pcen, range;
    void release() {
        itc.release();
        term.release();
        rows.release();
        X.release();
        P.release();
        FF.release();
        SF.release();
        dirj.release();
        pcen.release();
        range.release();
    }
};

namespace {

using namespace dev;

constexpr int kTrThreads = 256;
constexpr int kTrPairs = 8;     // pairs per chunk of k_tr_dirj
constexpr int kTrMaxFn = 16;    // radial functions of the training path (kTrPairs * kTrMaxFn threads fill g_n)
constexpr int kTrFeatY = 16;    // features per workgroup of k_tr_feat
constexpr int kTrMaxL = 12;     // angular momenta of the training path (kTrPairs * (kTrMaxL + 1) <= 128 threads)
constexpr int kTrL1 = kTrMaxL + 1;

struct TrArgs {
    const TrItc* itc;
    const TrTerm* term;
    const int32_t* rows;
    int n_des, n_coeff, n_gtinv;
    double* X;
    double* P;
    double* FF;
    double* SF;
    double* dirj;
    int32_t* pcen;
};

// a_{l,m} of factor `code` from the m <= 0 row value v
__device__ __forceinline__ void fac(int code, double& re, double& im) {
    if (code & (1 << 16)) {
        im = -im;
        if (code & (1 << 17)) {
            re = -re;
            im = -im;
        }
    }
}

__global__ __launch_bounds__(64) void k_tr_feat(DevModel M, DevSystem S, TrArgs T) {
    const int blk = blockIdx.x;
    const int lane = threadIdx.x;
    const int slot = blk * kBlock + lane;
    const int atom = S.atom_of_slot[slot];
    const int t = S.block_type[blk];
    const int f1 = min(T.n_des, (int)(blockIdx.y + 1) * kTrFeatY);
    for (int f = blockIdx.y * kTrFeatY; f < f1; ++f) {
        const int itc = f % T.n_gtinv, n = f / T.n_gtinv;
        const TrItc& I = T.itc[itc];
        bool seen = true;
        for (int mu = 0; mu < I.order; ++mu) seen = seen && I.spec[t][mu] >= 0;
        double d = 0.0;
        if (seen) {
            for (int k = I.tbeg; k < I.tend; ++k) {
                const TrTerm& tm = T.term[k];
                double pr = tm.c, pi = 0.0;
                for (int mu = 0; mu < I.order; ++mu) {
                    const int code = tm.code[mu];
                    const int lm = code & 0xffff;
                    double ar = S.A[a_index(M, blk, n, I.spec[t][mu], lm, 0, lane)];
                    double ai = S.A[a_index(M, blk, n, I.spec[t][mu], lm, 1, lane)];
                    fac(code, ar, ai);
                    const double nr = pr * ar - pi * ai;
                    pi = pr * ai + pi * ar;
                    pr = nr;
                }
                d += pr;
            }
        }
        if (atom >= 0) T.X[(size_t)atom * T.n_des + f] = d;
    }
}

__global__ __launch_bounds__(kTrThreads) void k_tr_poly(int natoms, TrArgs T) {
    const int64_t w = (int64_t)blockIdx.x * kTrThreads + threadIdx.x;
    if (w >= (int64_t)natoms * T.n_coeff) return;
    const int atom = (int)(w / T.n_coeff), p = (int)(w % T.n_coeff);
    double v = 1.0;
    for (int k = 0; k < 3; ++k) {
        const int f = T.rows[3 * p + k];
        if (f >= 0) v *= T.X[(size_t)atom * T.n_des + f];
    }
    T.P[w] = v;
}

__global__ __launch_bounds__(kTrThreads) void k_tr_pcen(DevSystem S, int32_t* pcen) {
    const int a = blockIdx.x * kTrThreads + threadIdx.x;
    if (a >= S.nown) return;
    for (int q = S.poff[a]; q < S.pend[a]; ++q) pcen[q] = a;
}

struct TrLds {
    double yt_diag[kTrL1 + 1];
    double yt_alm[kTrL1][kTrL1];
    double yt_ainv[kTrL1][kTrL1];
    double yt_grad[kTrL1][kTrL1];
    double ux[kTrPairs], uy[kTrPairs], uz[kTrPairs], rr[kTrPairs];
    int spec[kTrPairs], valid[kTrPairs];
    double gn[kTrPairs][kTrMaxFn], gd[kTrPairs][kTrMaxFn];
};

// dynamic LDS: a of the centre [n][s][lm] (re, im), then per chunk pair
// Y_{l,-m} and grad Y_{l,-m} [pair][lm][4] (re, im)
__global__ __launch_bounds__(kTrThreads) void k_tr_dirj(DevModel M, DevSystem S, TrArgs T) {
    extern __shared__ double tr_dyn[];
    __shared__ TrLds C;
    const int atom = blockIdx.x;
    if (atom >= S.nown) return;
    const int tid = threadIdx.x;
    const int nfn = M.n_fn, ns = M.n_type, nlmh = M.n_lmh, maxl = M.maxl;
    const int slot = S.slot_of_atom[atom];
    const int blk = slot / kBlock, lane = slot % kBlock;
    const int tcen = S.elem[atom];
    double* aL = tr_dyn;                                  // 2 * nfn * ns * nlmh
    double* yL = tr_dyn + (size_t)2 * nfn * ns * nlmh;    // kTrPairs * nlmh * 8
    for (int k = tid; k < nfn * ns * nlmh; k += kTrThreads) {
        const int lm = k % nlmh, s = (k / nlmh) % ns, n = k / (nlmh * ns);
        aL[2 * k] = S.A[a_index(M, blk, n, s, lm, 0, lane)];
        aL[2 * k + 1] = S.A[a_index(M, blk, n, s, lm, 1, lane)];
    }
    {
        const YTab<double>* yt = M.yt64;
        for (int k = tid; k < kTrL1 * kTrL1; k += kTrThreads) {
            const int l = k / kTrL1, mm = k % kTrL1;
            C.yt_alm[l][mm] = yt->alm[l][mm];
            C.yt_ainv[l][mm] = yt->ainv[l][mm];
            C.yt_grad[l][mm] = yt->grad[l][mm];
        }
        if (tid <= kTrL1) C.yt_diag[tid] = yt->diag[tid];
    }
    const int p0 = S.poff[atom], p1 = S.pend[atom];
    const double pir = 3.141592653589793 / M.cutoff;
    const int items = kTrPairs * T.n_gtinv * nfn;
    for (int pc = p0; pc < p1; pc += kTrPairs) {
        __syncthreads();
        if (tid < kTrPairs) {
            const int q = pc + tid;
            int ok = 0;
            if (q < p1) {
                const int nb = S.pnbr[q];
                double dx = S.x[3 * atom] - S.x[3 * nb], dy = S.x[3 * atom + 1] - S.x[3 * nb + 1],
                       dz = S.x[3 * atom + 2] - S.x[3 * nb + 2];
                if (S.pshift) {
                    dx += S.pshift[3 * q];
                    dy += S.pshift[3 * q + 1];
                    dz += S.pshift[3 * q + 2];
                }
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                ok = (r < M.cutoff) && (r > 0.0);
                if (ok) {
                    C.ux[tid] = dx / r;
                    C.uy[tid] = dy / r;
                    C.uz[tid] = dz / r;
                    C.rr[tid] = r;
                    C.spec[tid] = S.elem[nb];
                }
            }
            C.valid[tid] = ok;
        }
        __syncthreads();
        if (tid < kTrPairs * kTrL1) {
            // column m of pair j: Y_{l,-m} and grad Y_{l,-m} for l = m..maxl
            const int j = tid / kTrL1, m = tid % kTrL1;
            if (m <= maxl && C.valid[j]) {
                const double x = C.ux[j], y = C.uy[j], z = C.uz[j], ir = 1.0 / C.rr[j];
                double cr = 1.0, ci = 0.0;  // conj(u)^m
                for (int k = 0; k < m; ++k) {
                    const double nr = cr * x + ci * y;
                    ci = ci * x - cr * y;
                    cr = nr;
                }
                double mr = 0.0, mi = 0.0;  // conj(u)^(m-1)
                if (m > 0) {
                    mr = 1.0;
                    for (int k = 0; k < m - 1; ++k) {
                        const double nr = mr * x + mi * y;
                        mi = mi * x - mr * y;
                        mr = nr;
                    }
                }
                const double pr = cr * x + ci * y, pi = ci * x - cr * y;  // conj(u)^(m+1)
                const double sg = (m & 1) ? -1.0 : 1.0, zz = 1.0 - z * z, fm = (double)m;
                double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
                for (int l = m; l <= maxl; ++l) {
                    const double q = (l == m) ? C.yt_diag[m] : C.yt_alm[l][m] * (z * a1 - C.yt_ainv[l][m] * a2);
                    a2 = a1;
                    a1 = q;
                    double qp = 0.0;
                    if (l == m + 1)
                        qp = C.yt_diag[m + 1];
                    else if (l > m + 1)
                        qp = C.yt_alm[l][m + 1] * (z * b1 - C.yt_ainv[l][m + 1] * b2);
                    b2 = b1;
                    b1 = qp;
                    double gxr, gxi, gyr, gyi, gzr, gzi;
                    if (m == 0) {
                        const double c = C.yt_grad[l][0] * qp;
                        gxr = c * z * x;
                        gxi = 0.0;
                        gyr = c * z * y;
                        gyi = 0.0;
                        gzr = -c * zz;
                        gzi = 0.0;
                    } else {
                        const double c = C.yt_grad[l][m];
                        const double A = fm * z * q + zz * c * qp;
                        const double B = fm * q - z * c * qp;
                        const double E = fm * (1.0 + z * z) * q + z * zz * c * qp;
                        gxr = 0.5 * (-B * pr + E * mr);
                        gxi = 0.5 * (-B * pi + E * mi);
                        const double zr = B * pr + E * mr, zi = B * pi + E * mi;
                        gyr = 0.5 * zi;
                        gyi = -0.5 * zr;
                        gzr = -A * cr;
                        gzi = -A * ci;
                    }
                    double* o = yL + ((size_t)j * nlmh + (l * (l + 1) / 2 + l - m)) * 8;
                    o[0] = sg * q * cr;
                    o[1] = sg * q * ci;
                    o[2] = sg * gxr * ir;
                    o[3] = sg * gxi * ir;
                    o[4] = sg * gyr * ir;
                    o[5] = sg * gyi * ir;
                    o[6] = sg * gzr * ir;
                    o[7] = sg * gzi * ir;
                }
            }
        } else if (tid >= 128 && tid < 128 + kTrPairs * kTrMaxFn) {
            const int j = (tid - 128) / kTrMaxFn, n = (tid - 128) % kTrMaxFn;
            if (n < nfn && C.valid[j]) {
                const double r = C.rr[j];
                double sa, ca;
                sincos(pir * r, &sa, &ca);
                const double fc = 0.5 * (ca + 1.0), fcd = -0.5 * pir * sa;
                const double dd = r - M.mu[n], be = M.beta[n];
                const double b = exp(-be * (dd * dd));
                C.gn[j][n] = b * fc;
                C.gd[j][n] = b * (fcd - 2.0 * be * dd * fc);
            }
        }
        __syncthreads();
        for (int w = tid; w < items; w += kTrThreads) {
            const int n = w % nfn;
            const int itc = (w / nfn) % T.n_gtinv;
            const int j = w / (nfn * T.n_gtinv);
            const int q = pc + j;
            if (q >= p1) continue;
            const int f = n * T.n_gtinv + itc;
            double* out = T.dirj + ((size_t)q * T.n_des + f) * 3;
            double sy = 0.0, sx = 0.0, syy = 0.0, sz = 0.0;
            const TrItc& I = T.itc[itc];
            bool seen = C.valid[j] != 0;
            for (int mu = 0; mu < I.order; ++mu) seen = seen && I.spec[tcen][mu] >= 0;
            if (seen) {
                const int sj = C.spec[j];
                const double* ar = aL + (size_t)2 * n * ns * nlmh;
                const double* yj = yL + (size_t)j * nlmh * 8;
                for (int k = I.tbeg; k < I.tend; ++k) {
                    const TrTerm& tm = T.term[k];
                    for (int mu = 0; mu < I.order; ++mu) {
                        if (I.spec[tcen][mu] != sj) continue;
                        double pr = tm.c, pi = 0.0;
                        for (int mu2 = 0; mu2 < I.order; ++mu2) {
                            if (mu2 == mu) continue;
                            const int code = tm.code[mu2];
                            const double* av = ar + 2 * ((size_t)I.spec[tcen][mu2] * nlmh + (code & 0xffff));
                            double vr = av[0], vi = av[1];
                            fac(code, vr, vi);
                            const double nr = pr * vr - pi * vi;
                            pi = pr * vi + pi * vr;
                            pr = nr;
                        }
                        const int code = tm.code[mu];
                        const double* yv = yj + (size_t)(code & 0xffff) * 8;
                        double v[8];
                        for (int e = 0; e < 8; ++e) v[e] = yv[e];
                        for (int e = 0; e < 4; ++e) fac(code, v[2 * e], v[2 * e + 1]);
                        sy += pr * v[0] - pi * v[1];
                        sx += pr * v[2] - pi * v[3];
                        syy += pr * v[4] - pi * v[5];
                        sz += pr * v[6] - pi * v[7];
                    }
                }
            }
            if (seen) {
                const double gd = C.gd[j][n], gn = C.gn[j][n];
                out[0] = gd * C.ux[j] * sy + gn * sx;
                out[1] = gd * C.uy[j] * sy + gn * syy;
                out[2] = gd * C.uz[j] * sy + gn * sz;
            } else {
                out[0] = 0.0;
                out[1] = 0.0;
                out[2] = 0.0;
            }
        }
    }
}

// g_{q,p} = d P_cen(p) / d del_q
__device__ __forceinline__ void coef_grad(const TrArgs& T, int p, int cen, int64_t q, double g[3]) {
    const int f0 = T.rows[3 * p], f1 = T.rows[3 * p + 1], f2 = T.rows[3 * p + 2];
    const double* xc = T.X + (size_t)cen * T.n_des;
    const double* dq = T.dirj + (size_t)q * T.n_des * 3;
    double c0 = 1.0, c1 = 0.0, c2 = 0.0;
    if (f1 >= 0) {
        c0 = xc[f1];
        c1 = xc[f0];
        if (f2 >= 0) {
            c0 *= xc[f2];
            c1 *= xc[f2];
            c2 = xc[f0] * xc[f1];
        }
    }
    for (int k = 0; k < 3; ++k) {
        double v = c0 * dq[3 * f0 + k];
        if (f1 >= 0) v += c1 * dq[3 * f1 + k];
        if (f2 >= 0) v += c2 * dq[3 * f2 + k];
        g[k] = v;
    }
}

__global__ __launch_bounds__(kTrThreads) void k_tr_force(DevSystem S, TrArgs T, int pchunks) {
    const int atom = blockIdx.x / pchunks;
    const int p = (blockIdx.x % pchunks) * kTrThreads + threadIdx.x;
    if (atom >= S.nown || p >= T.n_coeff) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int e = S.goff[atom]; e < S.gend[atom]; ++e) {
        const int en = S.gent[e];
        const int64_t q = (en > 0 ? en : -en) - 1;
        double g[3];
        coef_grad(T, p, T.pcen[q], q, g);
        if (en > 0) {
            fx += g[0];
            fy += g[1];
            fz += g[2];
        } else {
            fx -= g[0];
            fy -= g[1];
            fz -= g[2];
        }
    }
    double* o = T.FF + ((size_t)atom * T.n_coeff + p) * 3;
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
}

__global__ __launch_bounds__(kTrThreads) void k_tr_stress(DevModel M, DevSystem S, TrArgs T, const int32_t* sbeg,
                                                         const int32_t* send) {
    __shared__ double red[6][kTrThreads];
    const int s = blockIdx.x / T.n_coeff, p = blockIdx.x % T.n_coeff;
    const int a0 = sbeg[s], a1 = send[s];
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (a1 > a0) {
        const int q0 = S.poff[a0], q1 = S.pend[a1 - 1];
        for (int q = q0 + threadIdx.x; q < q1; q += kTrThreads) {
            const int cen = T.pcen[q], nb = S.pnbr[q];
            double dx = S.x[3 * cen] - S.x[3 * nb], dy = S.x[3 * cen + 1] - S.x[3 * nb + 1],
                   dz = S.x[3 * cen + 2] - S.x[3 * nb + 2];
            if (S.pshift) {
                dx += S.pshift[3 * q];
                dy += S.pshift[3 * q + 1];
                dz += S.pshift[3 * q + 2];
            }
            double g[3];
            coef_grad(T, p, cen, q, g);
            // force term on the centre is -g
            v[0] -= dx * g[0];
            v[1] -= dy * g[1];
            v[2] -= dz * g[2];
            v[3] -= dy * g[2];
            v[4] -= dz * g[0];
            v[5] -= dx * g[1];
        }
    }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int h = kTrThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 6) T.SF[((size_t)s * T.n_coeff + p) * 6 + threadIdx.x] = red[threadIdx.x][0];
}

bool build_tables(mlip_ctx* c, TrainWork& w) {
    const Model& m = c->model;
    std::vector<TrItc> itc(m.n_gtinv);
    std::vector<TrTerm> term;
    std::vector<int32_t> tbeg(m.irreps.size()), tend(m.irreps.size());
    for (size_t r = 0; r < m.irreps.size(); ++r) {
        const Irrep& ir = m.irreps[r];
        tbeg[r] = static_cast<int32_t>(term.size());
        for (size_t k = 0; k < ir.m.size(); ++k) {
            TrTerm t{};
            t.c = ir.c[k];
            for (size_t mu = 0; mu < ir.l.size(); ++mu) {
                const int l = ir.l[mu], mm = ir.m[k][mu];
                const int am = mm < 0 ? -mm : mm;
                if (l > m.maxl) {
                    c->err = "invariant with l above the model's max_l";
                    return false;
                }
                int code = lmh_index(l, -am);
                if (mm > 0) code |= 1 << 16;
                if (mm > 0 && (mm & 1)) code |= 1 << 17;
                t.code[mu] = code;
            }
            term.push_back(t);
        }
        tend[r] = static_cast<int32_t>(term.size());
    }
    for (int i = 0; i < m.n_gtinv; ++i) {
        const Itc& it = m.itcs[i];
        TrItc& d = itc[i];
        std::memset(&d, 0, sizeof d);
        d.tbeg = tbeg[it.irrep];
        d.tend = tend[it.irrep];
        d.order = it.order;
        for (int t = 0; t < 2; ++t)
            for (int mu = 0; mu < 4; ++mu)
                d.spec[t][mu] = (t < m.n_type && mu < it.order) ? m.tcp_other[it.tc[mu]][t] : -1;
    }
    std::vector<int32_t> rows(3 * static_cast<size_t>(m.n_coeff), -1);
    for (int f = 0; f < m.n_des; ++f) rows[3 * f] = f;
    int p = m.n_des;
    for (const auto& cb : m.comb2) {
        rows[3 * p] = cb[0];
        rows[3 * p + 1] = cb[1];
        ++p;
    }
    for (const auto& cb : m.comb3) {
        rows[3 * p] = cb[0];
        rows[3 * p + 1] = cb[1];
        rows[3 * p + 2] = cb[2];
        ++p;
    }
    if (!upload(c, w.itc, itc.data(), itc.size(), "training tables") ||
        !upload(c, w.term, term.data(), term.size(), "training tables") ||
        !upload(c, w.rows, rows.data(), rows.size(), "training tables"))
        return false;
    w.built = true;
    return true;
}

}  // namespace

int train_features(mlip_ctx* c, double* X, double* P, double* FF, double* SF) {
    const Model& m = c->model;
    if (m.pair_model) {
        c->err = "training features are defined for the gtinv descriptor";
        return MLIP_ERR_MODEL;
    }
    static_assert(128 + kTrPairs * kTrMaxFn <= kTrThreads, "g_n threads of k_tr_dirj");
    static_assert(kTrPairs * kTrL1 <= 128, "Y_lm threads of k_tr_dirj");
    if (m.maxl > kTrMaxL) {
        c->err = "training features support max_l up to 12";
        return MLIP_ERR_MODEL;
    }
    if (m.n_type > 2 || m.n_fn > kTrMaxFn) {
        c->err = "training features support up to two species and 16 radial functions";
        return MLIP_ERR_MODEL;
    }
    for (const Itc& it : m.itcs)
        if (it.order > 4) {
            c->err = "training features support invariants up to order 4";
            return MLIP_ERR_MODEL;
        }
    if (!c->resident) {
        c->err = "no resident structure";
        return MLIP_ERR_STATE;
    }
    if (!c->trw) c->trw = new TrainWork();
    TrainWork& w = *c->trw;
    if (!w.built && !build_tables(c, w)) return MLIP_ERR_DEVICE;
    const DevSystem& s = c->ds;
    const int n = s.nown;
    const int64_t np = c->npairs;
    const size_t dbytes = static_cast<size_t>(np) * m.n_des * 3 * sizeof(double);
    const size_t fbytes = static_cast<size_t>(n) * m.n_coeff * 3 * sizeof(double);
    if (dbytes > (size_t(24) << 30) || fbytes > (size_t(24) << 30)) {
        c->err = "training features of this structure need more than 24 GB; use smaller batches";
        return MLIP_ERR_ARG;
    }
    const int nstruct = c->nstruct > 0 ? c->nstruct : 1;
    if (static_cast<int64_t>(nstruct) * m.n_coeff > INT_MAX) {
        c->err = "too many structures x coefficients";
        return MLIP_ERR_ARG;
    }
    if (!w.X.ensure(static_cast<size_t>(n) * m.n_des) || !w.P.ensure(static_cast<size_t>(n) * m.n_coeff) ||
        !w.FF.ensure(static_cast<size_t>(n) * m.n_coeff * 3) ||
        !w.SF.ensure(static_cast<size_t>(nstruct) * m.n_coeff * 6) ||
        !w.dirj.ensure(static_cast<size_t>(np) * m.n_des * 3) || !w.pcen.ensure(static_cast<size_t>(np))) {
        c->err = "device allocation failed for training features";
        return MLIP_ERR_DEVICE;
    }
    // single structure: one range over all atoms
    const int32_t* sb = c->sbeg;
    const int32_t* se = c->send;
    hipStream_t st = c->stream;
    if (c->nstruct == 0) {
        const int32_t r[2] = {0, n};
        if (!upload(c, w.range, r, 2, "structure range")) return MLIP_ERR_DEVICE;
        sb = w.range.p;
        se = w.range.p + 1;
    }
    TrArgs a{w.itc.p, w.term.p, w.rows.p, m.n_des, m.n_coeff, m.n_gtinv, w.X.p, w.P.p, w.FF.p, w.SF.p, w.dirj.p,
             w.pcen.p};
    if (!hip_ok(c, launch_anlm(c->dm, s, kFP64, st), "a_nlm launch")) return MLIP_ERR_DEVICE;
    if (s.nblocks > 0) {
        hipLaunchKernelGGL(k_tr_feat, dim3(s.nblocks, (m.n_des + kTrFeatY - 1) / kTrFeatY), dim3(64), 0, st, c->dm,
                           s, a);
        if (!hip_ok(c, launched(), "training feature launch")) return MLIP_ERR_DEVICE;
    }
    const int64_t npw = static_cast<int64_t>(n) * m.n_coeff;
    if (npw > 0) {
        hipLaunchKernelGGL(k_tr_poly, dim3(static_cast<unsigned>((npw + kTrThreads - 1) / kTrThreads)),
                           dim3(kTrThreads), 0, st, n, a);
        if (!hip_ok(c, launched(), "polynomial feature launch")) return MLIP_ERR_DEVICE;
    }
    if (n > 0 && np > 0 && (FF || SF)) {
        hipLaunchKernelGGL(k_tr_pcen, dim3((n + kTrThreads - 1) / kTrThreads), dim3(kTrThreads), 0, st, s, w.pcen.p);
        if (!hip_ok(c, launched(), "pair centre launch")) return MLIP_ERR_DEVICE;
        const size_t lds = sizeof(double) * (2 * static_cast<size_t>(m.n_fn) * m.n_type * m.n_lmh +
                                             static_cast<size_t>(kTrPairs) * m.n_lmh * 8);
        if (lds > lds_limit()) {
            c->err = "training features: the model's per-centre LDS tile exceeds the device limit";
            return MLIP_ERR_MODEL;
        }
        if (!hip_ok(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_tr_dirj),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)),
                    "training LDS"))
            return MLIP_ERR_DEVICE;
        hipLaunchKernelGGL(k_tr_dirj, dim3(n), dim3(kTrThreads), lds, st, c->dm, s, a);
        if (!hip_ok(c, launched(), "feature derivative launch")) return MLIP_ERR_DEVICE;
        const int pchunks = (m.n_coeff + kTrThreads - 1) / kTrThreads;
        if (FF) {
            hipLaunchKernelGGL(k_tr_force, dim3(static_cast<unsigned>(static_cast<int64_t>(n) * pchunks)),
                               dim3(kTrThreads), 0, st, s, a, pchunks);
            if (!hip_ok(c, launched(), "force feature launch")) return MLIP_ERR_DEVICE;
        }
        if (SF) {
            hipLaunchKernelGGL(k_tr_stress, dim3(static_cast<unsigned>(nstruct * m.n_coeff)), dim3(kTrThreads), 0, st,
                               c->dm, s, a, sb, se);
            if (!hip_ok(c, launched(), "stress feature launch")) return MLIP_ERR_DEVICE;
        }
    } else {
        if (FF && !hip_ok(c, hipMemsetAsync(w.FF.p, 0, fbytes, st), "force features")) return MLIP_ERR_DEVICE;
        if (SF && !hip_ok(c, hipMemsetAsync(w.SF.p, 0, sizeof(double) * nstruct * m.n_coeff * 6, st),
                          "stress features"))
            return MLIP_ERR_DEVICE;
    }
    if (!hip_ok(c, hipGetLastError(), "training kernels")) return MLIP_ERR_DEVICE;
    bool ok = true;
    if (X) ok = ok && hip_ok(c, hipMemcpyAsync(X, w.X.p, sizeof(double) * n * m.n_des, hipMemcpyDeviceToHost, st), "X");
    if (P) ok = ok && hip_ok(c, hipMemcpyAsync(P, w.P.p, sizeof(double) * npw, hipMemcpyDeviceToHost, st), "P");
    if (FF) ok = ok && hip_ok(c, hipMemcpyAsync(FF, w.FF.p, fbytes, hipMemcpyDeviceToHost, st), "FF");
    if (SF)
        ok = ok && hip_ok(c, hipMemcpyAsync(SF, w.SF.p, sizeof(double) * nstruct * m.n_coeff * 6,
                                            hipMemcpyDeviceToHost, st),
                          "SF");
    ok = ok && hip_ok(c, hipStreamSynchronize(st), "training features");
    return ok ? MLIP_OK : MLIP_ERR_DEVICE;
}

void train_free(mlip_ctx* c) {
    if (!c->trw) return;
    c->trw->release();
    delete c->trw;
    c->trw = nullptr;
}

}  // namespace mlip
