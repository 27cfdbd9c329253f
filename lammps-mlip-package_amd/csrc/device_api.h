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
};

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

void launch_anlm(const DevModel& m, const DevSystem& s, int precision, hipStream_t st);
void launch_q12(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st);
void launch_q34_adj(const DevModel& m, const DevSystem& s, hipStream_t st);
void launch_q34_feat(const DevModel& m, const DevSystem& s, hipStream_t st);
// general-loop passes for visits outside the unrolled signatures (launched
// by the two functions above when the model has such visits)
void launch_q34_dyn(const DevModel& m, const DevSystem& s, bool feat, hipStream_t st);
void launch_poly(const DevModel& m, const DevSystem& s, hipStream_t st);
// threads per workgroup of the force kernel: 64, 128 (default), 192 or 256
void launch_force(const DevModel& m, const DevSystem& s, int precision, int threads, hipStream_t st);
// per-structure energy + virial (7 doubles each) of a batch; soff = atom offsets
void launch_struct_reduce(const DevSystem& s, const int32_t* sbeg, const int32_t* send, int nstruct, double* out,
                          hipStream_t st);
void launch_gather(const DevModel& m, const DevSystem& s, hipStream_t st);
// radial-only style: features into D, energy and H = dE/dd, pair forces
void launch_pair_features(const DevModel& m, const DevSystem& s, hipStream_t st);
void launch_pair_poly(const DevModel& m, const DevSystem& s, hipStream_t st);
void launch_pair_force(const DevModel& m, const DevSystem& s, hipStream_t st);
// per-slot energies; with totals, also the fixed-order whole-system sums
void launch_energy_virial(const DevModel& m, const DevSystem& s, bool totals, hipStream_t st);
// plain streaming copy used to calibrate hardware byte counters
void launch_copy(void* dst, const void* src, int64_t nbytes, int reps, hipStream_t st);

}  // namespace mlip
