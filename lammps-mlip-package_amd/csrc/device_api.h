// Plain-data views of the model tables and of one structure's device
// buffers, passed by value to the HIP kernels, plus the launch wrappers the
// host engine calls.  All pointers are device pointers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "model.h"

namespace mlip {

constexpr int kBlock = kBlockLanes;  // atoms per slot block (one wavefront of lanes)


enum Precision : int { kFP64 = 0, kMixed = 1 };

// Constants of the normalised Legendre recursion (see pair_kernels.hip).
template <typename T>
struct YTab {
    T diag[kMaxL + 2];              // Qb_m^m
    T alm[kMaxL + 1][kMaxL + 1];    // recursion factor a_{l,m}; alm[m+1][m] = sqrt(2m+3)
    T ainv[kMaxL + 1][kMaxL + 1];   // 1 / a_{l-1,m} (0 for l < m + 2)
    T grad[kMaxL + 1][kMaxL + 1];   // sqrt((l+m+1)(l-m))
};

struct DevModel {
    int n_type, n_fn, n_gtinv, n_lmh, n_lmh3, maxl, lmax3, n_pitc, n_p;
    int has_q34;            // order-3/4 invariants present (k_q34 runs)
    int q12_fuse;           // order-1/2 adjoint rows L <= q12_fuse are the order-3/4 kernel's (-1: none)
    int q12_fuse_hi;        // ... and the rows lmax3 < L <= q12_fuse_hi (a_nlm from HBM)
    int q12_adj_levels;     // 1 if some level is left to the order-1/2 adjoint kernel (else it
                            // is not launched and its energies Eq12 are not read)
    double cutoff;
    double beta[kMaxFnPair], mu[kMaxFnPair];
    const double* coef;     // dense coefficient blocks
    const double* weight;   // coeff/scale, linear part first (n_des)
    const VisitQ12* q12;    // all types
    const int32_t* q12_off; // [type][maxl+2]
    const VisitPack* vadj;
    const GroupQ34* gadj;
    int32_t gadj_beg[2], gadj_end[2];
    const VisitPack* vfeat;
    const GroupQ34* gfeat;
    int32_t gfeat_beg[2], gfeat_end[2];
    // polynomial-feature slots (pslot) no feature visit of a species writes:
    // zlist[zoff[t] .. zoff[t+1]) are set to zero by the order-1/2 feature pass
    const int32_t* zlist;
    int32_t zoff[3];
    // order-1/2 polynomial features per centre species, fq12[fq12_off[t] ..
    // fq12_off[t+1]) (formed in the a_nlm kernel's epilogue)
    const FeatQ12* fq12;
    int32_t fq12_off[3];
    // Polynomial features on the device are per species and compact: a
    // centre of species t keeps the ranks [lo, hi) = prange[t] it sees, as
    // the row [n][rank - lo] of n_fn * (hi - lo) doubles in a slot row of
    // n_pr (the largest over species); lin_p ([t][n_pr]) and c3_off
    // ([t][n_pr + 1]) use the same per-species columns.
    int32_t prange[2][2];
    int32_t n_pr;
    const double* lin_p;    // linear weights of polynomial features ([species][n_pr]; n_p for mlip_pair)
    const int32_t* c3_off;  // cubic terms transposed by target: offsets ([species][n_pr + 1])
    const int32_t* c3_bc;   // 2 per entry (columns of the same species)
    const double* c3_w;
    // 1 where the model has no cubic terms: the quadratic-form GEMM's
    // epilogue forms H = lin + D W and per column tile partial energies
    // sum d (lin + (D W) / 2), which launch_poly sums into Ep
    int poly_fused;
    // per-atom energy split of the reference CPU style (each polynomial term
    // of dE/da weighted by 1 / its degree in a): linear weights divided by
    // the family's order (n_des), the degree (family order) of each compact
    // polynomial column ([species][n_pr]) and the per-species quadratic forms
    const double* weight_e;
    const int32_t* pdeg;
    const double* wmat;
    int32_t wofs[2], wdim[2];
    // radial-only style
    int pair_model;         // 1: features sum_j g_n(r) per type combination
    int pair_tc[2][2];      // [centre type][neighbor type] -> type combination
    const double* pair_w;   // linear weights, n_p (= n_des) entries
    const int32_t* ph_off;  // per-feature derivative lists
    const PairHEntry* ph_ent;
    int q34_red;            // doubles of LDS partial rows of split adjoint groups
    int q34_dyn_adj, q34_dyn_feat;  // visits for the general-loop order-3/4 kernel
    const YTab<double>* yt64;
    const YTab<float>* yt32;
    // orders 5/6: visits of every species (adjoint, then features), their
    // term codes and coefficients; tile of a_nlm up to l = lmaxh
    const VisitHi* hvis;
    const int32_t* hterm;
    const double* hcoef;
    int32_t hadj_beg[2], hadj_end[2], hfeat_beg[2], hfeat_end[2];
    int lmaxh, n_lmhh;
};

// order-1/2 adjoint row L runs inside the order-3/4 kernel
__host__ __device__ inline bool q12_fused(const DevModel& m, int L) {
    return L <= m.q12_fuse || (L > m.lmax3 && L <= m.q12_fuse_hi);
}

struct DevSystem {
    int nall;      // positions (owned atoms first; ghosts / images after)
    int nown;      // owned atoms (centres)
    int nblocks;   // slot blocks of kBlock same-type centres
    int64_t npairs;
    int max_neigh;
    const double* x;          // [nall][3]
    const int32_t* elem;      // [nall] element index
    const int32_t* poff;      // [nown] first pair of each centre
    const int32_t* pend;      // [nown] end of each centre's pairs (poff + 1 for one CSR)
    const int32_t* pnbr;      // [npairs] neighbor position index
    const int32_t* pcen;      // [npairs] centre position index, or null (= owner)
    const double* pshift;     // [npairs][3] added to x_c - x_n, or null
    const int32_t* atom_of_slot;  // [nblocks*kBlock], -1 for padding
    const int32_t* slot_of_atom;  // [nown]
    const int32_t* block_type;    // [nblocks]
    double* A;       // a_nlm, block-interleaved [blk][n][s][lm][re,im][64]
    double* G;       // adjoint dE/da, same layout
    double* D;       // polynomial features [slot][n_p]
    double* Hq;      // quadratic-form product [slot][n_p]
    double* H;       // polynomial adjoint [slot][n_p]
    double* Ep;      // [slot]
    double* Eq12;    // [slot][n_fn]
    double* Eq34;    // [slot][n_fn][kWavesPerWg]
    double* gpair;   // [npairs][3] dE_centre/d(del)
    double* vown;    // [nown][6]
    const int32_t* goff;   // [nall] force gather lists: begin
    const int32_t* gend;   // [nall] end (goff + 1 for one CSR)
    const int32_t* gent;   // signed pair entries: p+1 (neighbor side) / -(p+1) (centre side)
    double* f;       // [nall][3]
    double* e_slot;  // [nblocks*kBlock]
    double* red;     // scratch for reductions
    double* out;     // [7]: energy, virial[6]
};

// Host-side construction of the recursion tables (uploaded per context).
void build_ylm_tables(YTab<double>& t64, YTab<float>& t32);

// Launch wrappers return the first HIP error of their launches (a refused
// launch configuration, an LDS request beyond the device's limit, an invalid
// kernel argument), hipSuccess otherwise; the engine turns any error into a
// failed call.  They enqueue work only: execution errors surface at the next
// synchronisation.
inline hipError_t first_error(hipError_t a, hipError_t b) { return a != hipSuccess ? a : b; }
// Largest dynamic + static LDS one workgroup may use on the current device (bytes).
size_t lds_limit();
// hipGetLastError after a launch, with a dynamic LDS request checked first.
hipError_t launched(size_t lds_bytes = 0);

// a_nlm kernels: kAnlmValu (k_anlm, thread per lm column, FP64 registers) or
// kAnlmMfma (k_anlm_mfma, chunk products on the matrix cores) where the
// model fits it (anlm_has_mfma), k_anlm otherwise
enum AnlmKind : int { kAnlmValu = 0, kAnlmMfma = 1 };
// feat: the register kernel also forms the order-1/2 polynomial features
// (what k_q12<true> does) from the a_nlm it holds; only where
// anlm_forms_features says the launch can
hipError_t launch_anlm(const DevModel& m, const DevSystem& s, int precision, hipStream_t st, int kind = kAnlmValu,
                       bool feat = false);
bool anlm_has_mfma(const DevModel& m);
bool anlm_forms_features(const DevModel& m, int precision, int kind);
hipError_t launch_q12(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st);
hipError_t launch_q34_adj(const DevModel& m, const DevSystem& s, hipStream_t st);
hipError_t launch_q34_feat(const DevModel& m, const DevSystem& s, hipStream_t st);
// general-loop passes for visits outside the unrolled signatures (launched
// by the two functions above when the model has such visits)
hipError_t launch_q34_dyn(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st);
hipError_t launch_poly(const DevModel& m, const DevSystem& s, hipStream_t st);
// orders 5/6: polynomial features (feat) or their share of the adjoint G and
// of the energy of linear families (qhigh.hip); no-op for models without them
hipError_t launch_qhigh(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st);
// Hq = D W per species on the matrix cores (slot blocks grouped by species,
// nblk_type[t] blocks of species t in order); for poly_fused models H =
// lin + D W and per column tile partial energies (in the Hq rows) instead
hipError_t launch_poly_gemm(const DevModel& m, const DevSystem& s, const int* nblk_type, hipStream_t st);
// Ep from the fused GEMM epilogue's partial energies (poly_fused models)
hipError_t launch_poly_sum(const DevModel& m, const DevSystem& s, hipStream_t st);
// threads per workgroup of the force kernel: 64, 128, 192 (default) or 256
hipError_t launch_force(const DevModel& m, const DevSystem& s, int precision, int threads, hipStream_t st);
// per-structure energy + virial (7 doubles each) of a batch; soff = atom offsets
hipError_t launch_struct_reduce(const DevSystem& s, const int32_t* sbeg, const int32_t* send, int nstruct,
                                double* out, hipStream_t st);
hipError_t launch_gather(const DevModel& m, const DevSystem& s, hipStream_t st);
// radial-only style: features into D, energy and H = dE/dd, pair forces
hipError_t launch_pair_features(const DevModel& m, const DevSystem& s, hipStream_t st);
hipError_t launch_pair_poly(const DevModel& m, const DevSystem& s, hipStream_t st);
hipError_t launch_pair_force(const DevModel& m, const DevSystem& s, hipStream_t st);
// per-slot energies; with totals, also the fixed-order whole-system sums
hipError_t launch_energy_virial(const DevModel& m, const DevSystem& s, bool totals, hipStream_t st);
// Per-atom energy and virial in the reference CPU style's split
// (pair_mlip_gtinv.cpp:240-253 with ev_tally_xyz; pair_mlip_pair.cpp:166-230):
// launch_poly_split writes He = dE_i/dd with every polynomial term weighted
// by 1 / its degree in a (into s.Hq); the adjoint kernels run on that He and
// weight_e give the split adjoint in s.G; launch_pair_energy writes, per
// directed pair p, 7 doubles: its energy share sum w_m Re(G phi_p) and
// del_p (x) (-g_p) from s.gpair (xx, yy, zz, xy, xz, yz); launch_atom_ev
// sums half of them over each position's gather list (evat[nall][7]).
hipError_t launch_poly_split(const DevModel& m, const DevSystem& s, hipStream_t st);
hipError_t launch_pair_energy(const DevModel& m, const DevSystem& s, double* epv, hipStream_t st);
hipError_t launch_pair_model_energy(const DevModel& m, const DevSystem& s, double* epv, hipStream_t st);
hipError_t launch_atom_ev(const DevSystem& s, const double* epv, double* evat, hipStream_t st);
// n doubles of dst set to zero by a kernel (no memset node in a captured graph)
hipError_t launch_fill(double* dst, size_t n, hipStream_t st);
// plain streaming copy used to calibrate hardware byte counters
hipError_t launch_copy(void* dst, const void* src, int64_t nbytes, int reps, hipStream_t st);

}  // namespace mlip
