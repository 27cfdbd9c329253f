// Host-side model layer of the gtinv evaluation path: potential-file parser,
// generalized Clebsch-Gordan table, feature / polynomial index enumeration
// and the flattened per-species tables the HIP kernels consume.
//
// Reference counterparts:
//   potential file + initialize     src/USER-MLIP/mlip_model.h:35-120
//   gtinv screening                 src/USER-MLIP/mlip_read_gtinv.cpp:35-76
//   (irreps x type-comb) features   src/USER-MLIP/mlip_model_params.cpp:86-126
//   polynomial index enumeration    src/USER-MLIP/mlip_model_params.cpp:27-82,128-170
//   model_type 1 products           src/USER-MLIP/mlip_polynomial.cpp:36-40
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace mlip {

constexpr int kMaxFn = 32;      // radial functions supported by the kernels
constexpr int kMaxL = 20;       // max angular momentum of a_nlm (the gtinv table's largest l)
constexpr int kMaxOrderGpu = 6; // highest invariant order of the HIP path (the table's largest)
constexpr int kMaxLQ12 = 12;    // order-1/2 levels up to this l run with the row in registers
constexpr int kWavesPerWg = 12; // waves sharing one a_nlm tile in the order-3/4 kernel
constexpr int kBlockLanes = 64; // centres per slot block (the lanes of one wave)
constexpr int kUnrollCap3 = 8;  // order-3 visits with all l <= 8 run fully unrolled (q34_signatures.inc)
constexpr int kUnrollCap4 = 2;  // order-4 visits with all l <= 2 run fully unrolled

struct Irrep {
    std::vector<int> l;                // nondecreasing l_1..l_q
    std::vector<std::vector<int>> m;   // m-tuples with sum 0
    std::vector<double> c;             // coefficients (reference decimal strings)
};

struct GtinvTable {
    std::vector<Irrep> irreps;
};

// The gtinv table is compiled into the library (see gtinv_blob.S).
bool load_builtin_gtinv(GtinvTable& out, std::string& err);
bool parse_gtinv_text(const char* text, size_t len, GtinvTable& out, std::string& err);

// One structural feature family: an irrep combined with a multiset of type
// combinations (LinearTermGtinv, mlip_model_params.h:88-92).
struct Itc {
    int irrep;            // index into Model::irreps (screened)
    int order;            // q
    std::array<int, 6> tc;
    unsigned type_mask;   // bit t set <=> centre type t sees this feature
};

// ---- device table records (plain data, copied to HBM) ----
// Order-1/2 visit (cheap orders, evaluated from HBM-resident a_nlm).
struct VisitQ12 {
    int32_t order;     // 1 or 2
    int32_t itc;
    int32_t s0, s1;    // species of the two factors (s1 unused for order 1)
    int32_t pslot;     // rank among polynomial itcs or -1
    int32_t coef;      // offset into coefficient array: order 2 -> 2L+1 values (m0=-L..L)
};

// Order-1/2 polynomial feature of one centre species, formed in the a_nlm
// kernel's epilogue: d = C a_{s0,0,0} (order 1) or
// sum_{m0} C(m0) a_{s0,L,m0} a_{s1,L,-m0} (order 2), stored at rank offset
// col (rank - lo of the species) of each radial index's row segment.
struct FeatQ12 {
    int32_t order, L, s0, s1, col, coef;
};

// Order-3/4 visit: contributions of one (itc, factor position) to the
// adjoint G_{S,L,m<=0} and, for the designated position, the feature d.
struct VisitQ34 {
    int32_t order;     // 3 or 4; 1 or 2 for fused order-1/2 adjoint rows; 0: empty row marker
    int32_t itc;
    int32_t la, sa;    // other factors in irrep order
    int32_t lb, sb;
    int32_t lc, sc;    // order 4 only
    int32_t pslot;     // rank among polynomial itcs or -1
    int32_t flags;     // bit0: compute d here; bit1: accumulate G; bit2: unrolled signature
    int32_t coef;      // dense coefficient block offset
    int32_t dcoef;     // dense pair-summed block offset (flags bit0)
    // species-paired adjoint groups (GroupQ34::S == kPairedRows): the rows
    // the visit adds into (0: G_{0,L}, 1: G_{1,L}, 2: both, with the
    // coefficient block both families share); pslot1 / itc1 are the second
    // family's H source when row == 2
    int32_t row;
    int32_t pslot1, itc1;
};

// Device form of a VisitQ34 (16 bytes, one scalar load): sig packs order,
// l_a, l_b, l_c, s_a, s_b, s_c and flags in 4 bits each (low to high); the
// feature-pass visits read their coefficients at coef (dcoef == coef there)
// and keep pslot / itc.  Adjoint visits carry the H source of their family
// as h = pslot >= 0 ? pslot : -1 - itc in the pslot field (and the second
// family's in the itc field for a visit of both paired rows); flags bit0
// marks a visit of row 1 of a paired group, bit3 one of both rows.
struct VisitPack {
    int32_t sig;
    int32_t coef;
    int32_t pslot;
    int32_t itc;
};

// Output group: all visits of one (species S, angular momentum L) handled by
// one wave of the adjoint kernel with a statically unrolled G_L accumulator.
// Output-stationary group of the order-3/4 kernel: visits that all add into
// G_{S,L,m}.  A heavy (S, L) row is split into nparts sub-groups on
// different waves; each writes its partial row to LDS at `red` (doubles
// from the partial area) and the wave of part 0 sums them in part order.
// S == kPairedRows: rows (0, L) and (1, L) of a two-species model handled
// together, so that the visits both rows have in common (same factors, same
// coefficients, different families) form their products once.
constexpr int kPairedRows = 2;
// largest L whose two species rows are paired (the kernel holds both rows and
// a visit accumulator, 6 (L + 1) doubles): at 256k atoms of the order-4
// benchmark model the order-3/4 kernel takes 38.1 ms unpaired, 36.8 / 35.1 /
// 38.6 ms with rows up to L = 4 / 5 / 6 paired (at 6 the registers spill)
constexpr int kPairMaxL = 5;
struct GroupQ34 {
    int32_t S, L;
    int32_t vbeg, vend;
    int32_t wave;      // wave of the workgroup that owns the group
    int32_t part;      // sub-group index within its (S, L) row
    int32_t nparts;    // sub-groups of the row (1: not split)
    int32_t red;       // partial-row offset of this sub-group (split rows only)
};

// Order-5/6 visit (the table's order-5 and order-6 invariants, all factors
// l <= 1): one (feature family, factor position) with its terms listed
// explicitly.  Term codes hold k (output m = -k) in bits 0-3 and m + 8 of
// each other factor in 4 bits from bit 4; coefficients are parallel doubles.
// Adjoint visits add H C prod(other factors) into G_{S,L,-k}; feature visits
// (position 0 of a polynomial family) form d = sum C' Re(a_{S,L,-k} prod)
// with C' the coefficient summed with that of the negated m-tuple for k > 0.
constexpr int kMaxLHi = 4;      // largest l of an order-5/6 factor the kernel holds
struct VisitHi {
    int32_t S, L;
    int32_t order, itc, pslot;
    int32_t ls[5];              // other factors: l | species << 8
    int32_t tbeg, tend;         // term range
};

// Radial-only model (des_type pair): one entry of the per-feature
// derivative lists, H_a = w_a + sum_entries w prod(d_b, d_c) (b, c = -1 for
// absent factors); ew = w / order is the Euler energy share of the term.
struct PairHEntry {
    int32_t b, c;
    double w, ew;
};

constexpr int kMaxFnPair = 32;  // radial functions of the radial-only style

// An order-3/4 visit before grouping: output row (S, L), the visit, its
// term count and the length of its coefficient block.
struct PlanVisit {
    int S, L;
    VisitQ34 v;
    long cost;
    int clen;
};

struct TypeTables {
    // order 1/2: per L, range into visits_q12
    std::vector<int32_t> q12_off;      // size maxl+2
    std::vector<VisitQ12> visits_q12;
    // order 3/4: adjoint pass and polynomial-feature pass
    std::vector<GroupQ34> groups_adj, groups_feat;
    std::vector<VisitQ34> visits_adj, visits_feat;
    // orders 5/6: adjoint and polynomial-feature visits
    std::vector<VisitHi> hi_adj, hi_feat;
    // order 3/4 visits before grouping (plan_q34 groups them)
    std::vector<PlanVisit> plan_adj, plan_feat;
};

struct Model {
    // file contents
    std::vector<std::string> elements;
    std::vector<double> masses;
    double cutoff = 0.0;
    int model_type = 0, maxp = 0, maxl = 0, gtinv_order = 0;
    std::vector<int> gtinv_maxl, gtinv_sym;
    std::vector<double> beta, mu;
    std::vector<double> coeff_file, scale_file;
    // enumeration
    int n_type = 0, n_fn = 0, n_tc = 0, n_gtinv = 0, n_des = 0, n_coeff = 0;
    std::vector<Irrep> irreps;
    std::vector<Itc> itcs;
    std::vector<std::array<int, 2>> tcp_other;   // [tc][type] -> species or -1
    std::vector<double> weight;                  // coeff/scale, size n_coeff
    std::vector<int> poly_features;              // sorted feature ids in products
    std::vector<std::array<int, 2>> comb2;
    std::vector<std::array<int, 3>> comb3;
    // GPU-side layout
    int lmax3 = 0;             // largest l among order>=3 irreps (LDS tile)
    int n_lmh = 0;             // (maxl+1)(maxl+2)/2, m<=0 entries of a_nlm
    int n_lmh3 = 0;            // same for lmax3
    int n_pitc = 0;            // itcs whose features enter polynomial products
    std::vector<int> prank;    // itc -> rank among polynomial itcs or -1
    int n_p = 0;               // n_fn * n_pitc; feature (n, rank) = n * n_pitc + rank
    int prange[2][2] = {{0, 0}, {0, 0}};  // ranks a centre species sees: [lo, hi)
    std::vector<double> wmat;  // n_p x n_p symmetric quadratic form
    std::vector<int32_t> c3_idx;  // comb3 in polynomial-feature space, 3 per term
    std::vector<double> c3_w;
    std::vector<double> lin_p;    // linear weights of the polynomial features (n_p)
    std::vector<double> coef;     // dense coefficient blocks for all visits
    std::vector<TypeTables> types;
    bool has_order34 = false;
    bool has_feat34 = false;
    // order-1/2 adjoint rows L <= q12_fuse run inside the order-3/4 adjoint
    // kernel, on its LDS tile (-1: none; the order-1/2 kernel takes the rest)
    int q12_fuse = -1;
    // ... and so do the rows above the order-3/4 tile, lmax3 < L <= q12_fuse_hi
    // (their a_nlm read from HBM)
    int q12_fuse_hi = -1;
    // radial-only style (des_type pair): features f = tc * n_fn + n
    bool pair_model = false;
    std::array<std::array<int, 2>, 2> pair_tc{};   // [centre type][neighbor type] -> tc
    std::vector<int32_t> ph_off;                   // n_des + 1
    std::vector<PairHEntry> ph_ent;
    int q34_red = 0;           // doubles of LDS partial rows (adjoint pass), max over types
    // orders 5/6: term codes and coefficients of every VisitHi, tile l_max
    std::vector<int32_t> hi_term;
    std::vector<double> hi_coef;
    int lmaxh = -1;            // largest l of an order-5/6 factor (-1: none)
    bool has_hi = false;
    bool q34_dyn_adj = false;  // visits outside the unrolled signatures (general loop kernel)
    bool q34_dyn_feat = false;
    // work model of the order-3/4 workgroup plan (terms per scheduling
    // region and per visit); 3 and 24 were the best of a sweep at the 256k
    // benchmark model, mlip_autotune_q34 times alternatives on the device
    int plan_region = 3, plan_visit = 24;
    int plan_pair = kPairMaxL;  // largest L of species-paired rows (-1: none)
};

// (Re)group the order-3/4 visits of every centre type over the waves of the
// order-3/4 workgroup with the given work model.
void plan_q34(Model& m, int region_cost, int visit_cost, int pair_lmax);

// Parse a potential file (format of mlip_model.h:45-101) and build every
// table.  Returns false with a message instead of exiting.
bool load_model(const std::string& path, const GtinvTable& table, Model& m, std::string& err,
                bool strict_count = true);

inline int lmh_index(int l, int m) { return l * (l + 1) / 2 + l + m; }  // m <= 0
inline int n_lmh_of(int lmax) { return (lmax + 1) * (lmax + 2) / 2; }

}  // namespace mlip
