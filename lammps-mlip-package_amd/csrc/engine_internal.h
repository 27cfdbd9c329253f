// Internal types of the host engine shared by its translation units
// (engine.hip, pool.hip): device buffers, the context behind the C ABI and
// the pipeline entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mlip_hip.h"
#include "device_api.h"
#include "model.h"

namespace mlip {

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    bool ensure(size_t n) {
        if (n <= cap && p) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = n;
        return true;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct ProfEvent {
    const char* name;
    hipEvent_t a, b;
};

struct Pool;    // device-resident request pool (pool.hip)
struct NbWork;  // device neighbor-build scratch (nbuild.hip)
struct TrainWork;  // training-feature tables and outputs (train.hip)

}  // namespace mlip

struct mlip_ctx {
    mlip::Model model;
    std::vector<int> elem_of_type;  // LAMMPS type-1 -> element (or -1)
    int precision = mlip::kFP64;
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;
    // model tables
    mlip::DBuf<double> d_coef, d_weight, d_wmat, d_linp, d_c3w, d_weight_e;
    mlip::DBuf<mlip::VisitQ12> d_q12;
    mlip::DBuf<mlip::FeatQ12> d_fq12;
    mlip::DBuf<int32_t> d_q12off, d_c3off, d_c3bc, d_pdeg;
    mlip::DBuf<mlip::VisitPack> d_vadj, d_vfeat;
    mlip::DBuf<mlip::GroupQ34> d_gadj, d_gfeat;
    mlip::DBuf<mlip::YTab<double>> d_yt64;
    mlip::DBuf<mlip::YTab<float>> d_yt32;
    mlip::DBuf<double> d_pairw;
    mlip::DBuf<int32_t> d_phoff, d_zlist;
    mlip::DBuf<mlip::PairHEntry> d_phent;
    mlip::DBuf<mlip::VisitHi> d_hvis;
    mlip::DBuf<int32_t> d_hterm;
    mlip::DBuf<double> d_hcoef;
    mlip::DevModel dm{};
    // structure buffers
    mlip::DBuf<double> x, pshift, A, G, D, Hq, H, Ep, Eq12, Eq34, gpair, vown, f, e_slot, red, out;
    // per-atom energy and virial in the reference's pair split: per directed
    // pair (energy share, virial[6]) and per position (the same 7 sums)
    mlip::DBuf<double> epv, evat;
    int atom_mode = MLIP_ATOM_PAIR_SPLIT;
    mlip::DBuf<int32_t> elem, poff, pnbr, pcen, atom_of_slot, slot_of_atom, block_type, goff, gent;
    mlip::DevSystem ds{};
    bool resident = false;
    int resident_natoms = 0;
    // profiling: events recorded around every launch while enabled, folded
    // into per-name totals at each synchronisation
    bool prof = false;
    std::vector<mlip::ProfEvent> events;
    size_t ev_used = 0;
    std::vector<std::string> prof_names;
    std::vector<double> prof_ms;
    std::vector<int> prof_count;
    // algorithmic work of the current structure (see mlip_flops)
    std::vector<int64_t> owners_per_type;
    int64_t npairs = 0;
    // launch shape chosen by mlip_autotune
    int force_threads = 192;
    int anlm_kind = mlip::kAnlmMfma;  // a_nlm kernel (mlip_autotune keeps the faster one)
    int nblk_type[2] = {0, 0};  // slot blocks of each species (blocks are grouped by species)
    size_t wofs[2] = {0, 0};    // per-species quadratic form in d_wmat: offset and dimension
    int wdim[2] = {0, 0};
    // batch of independent structures (0 = single structure): per-structure
    // atom ranges [sbeg, send) and 7 results each
    int nstruct = 0;
    mlip::DBuf<int32_t> soff;
    const int32_t* sbeg = nullptr;
    const int32_t* send = nullptr;
    mlip::DBuf<double> sout;
    double* sres = nullptr;  // where the per-structure results go
    // whole-system energy/virial reductions (off for the request pool)
    bool total_reduce = true;
    // one evaluation captured as a HIP graph and replayed
    bool use_graph = false;
    bool graph_warm = false;  // one eager evaluation ran since the last change
    hipGraphExec_t gexec = nullptr;
    // device-resident request pool, when initialised
    mlip::Pool* pool = nullptr;
    // neighbor lists of mlip_batch_set: built on the device (default) or on
    // host threads; scratch of the device builder
    bool nb_device = true;
    mlip::NbWork* nbw = nullptr;
    // training features (mlip_train_features), allocated on first use
    mlip::TrainWork* trw = nullptr;
};

namespace mlip {

bool hip_ok(mlip_ctx* c, hipError_t e, const char* what);

// Device buffers of one structure (positions, lists, slots, work arrays).
struct SysBufs {
    double *x, *pshift, *A, *G, *D, *Hq, *H, *Ep, *Eq12, *Eq34, *gpair, *vown, *f, *e_slot, *red, *out;
    int32_t *elem, *poff, *pnbr, *pcen, *goff, *gent, *atom_of_slot, *slot_of_atom, *block_type;
};
// Element counts of the slot-indexed work arrays of a structure.
struct WorkSizes {
    size_t nblocks, nslot, asz, psz;  // asz: a_nlm / G doubles, psz: D / Hq / H doubles
};
WorkSizes work_sizes(const Model& m, int nall, int nown, int64_t npairs, int nblocks);
void assign_slots(const mlip_ctx* c, int nown, const std::vector<int>& elem_of_owner,
                  std::vector<int32_t>& atom_of_slot, std::vector<int32_t>& slot_of_atom,
                  std::vector<int32_t>& block_type);

// Slot arrays uploaded into b, a_nlm and the per-centre virial cleared,
// c->ds pointed at b (positions, elements and lists must be in b already).
bool bind_system(mlip_ctx* c, int nall, int nown, int64_t npairs, const std::vector<int32_t>& elem_own,
                 const std::vector<int32_t>& atom_of_slot, const std::vector<int32_t>& slot_of_atom,
                 const std::vector<int32_t>& block_type, const SysBufs& b);

template <typename T>
bool upload(mlip_ctx* c, DBuf<T>& b, const T* h, size_t n, const char* what) {
    if (!b.ensure(n)) {
        c->err = std::string("device allocation failed for ") + what;
        return false;
    }
    if (n == 0) return true;
    return hip_ok(c, hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream), what);
}

void drop_graph(mlip_ctx* c);
bool run_pipeline(mlip_ctx* c);
bool run_eval(mlip_ctx* c);
bool run_atom_split(mlip_ctx* c, double* epv, double* evat);
// LAMMPS-shaped input (mlip_compute / mlip_compute_flat) turned into the
// engine's directed-pair lists: both directions of every pair, centre
// positions kept (pcen), owners in ilist order, force gather lists.
struct LammpsLists {
    int nall = 0;
    int64_t npairs = 0;
    std::vector<double> x;
    std::vector<int32_t> elem, elem_own, poff, pnbr, pcen, gof, gen;
};
int lammps_lists(mlip_ctx* c, int nlocal, int nghost, const double* x, const int* type, const int64_t* tag, int inum,
                 const int* ilist, const int* numneigh, const int* const* firstneigh, const int* neigh_off,
                 const int* neigh_idx, int neigh_full, LammpsLists& L);
// After an evaluation of such lists: adds forces, energy and virial into the
// caller's arrays and fills eatom / vatom (epv, evat: device scratch of
// 7 npairs and 7 nall doubles for the pair split, unused in site mode).
int lammps_results(mlip_ctx* c, int nall, int inum, const int* ilist, double* f, double* eng, double* virial,
                   double* eatom, double* vatom, double* epv, double* evat);
void collect_profile(mlip_ctx* c);
void build_gather(int nall, const std::vector<int32_t>& poff, const std::vector<int32_t>& pnbr,
                  const std::vector<int32_t>* pcen, std::vector<int32_t>& goff, std::vector<int32_t>& gent);
void pool_free(mlip_ctx* c);
// Full periodic lists (0 < r < rc, self-images, image shifts) of nstruct
// structures whose positions (concatenated, natoms[s] each) are already in
// c->x, built on the device into c->poff, c->pnbr, c->pshift, with the force
// gather lists in c->goff, c->gent; identical to build_full_list followed by
// build_gather.  Synchronises once to size the pair arrays.
bool device_lists(mlip_ctx* c, int nstruct, const int32_t* natoms, const double* cells, double rc, int64_t& npairs);
void nb_free(mlip_ctx* c);
// The same build with explicit outputs, a cutoff per structure and the
// work on stream st: positions (concatenated) in x; lists into poff, pnbr,
// pshift, goff, gent; spairs (nstruct + 1, optional) receives the first pair
// of each structure.  Synchronises st once.
struct NbOut {
    DBuf<int32_t>* poff;
    DBuf<int32_t>* pnbr;
    DBuf<double>* pshift;
    DBuf<int32_t>* goff;
    DBuf<int32_t>* gent;
};
bool device_lists_into(mlip_ctx* c, NbWork*& w, hipStream_t st, int nstruct, const int32_t* natoms,
                       const double* cells, const double* rcs, const double* x, const NbOut& out, int64_t& npairs,
                       std::vector<int64_t>* spairs);
void nb_work_free(NbWork*& w);
// Training outputs of the resident structure or batch (see
// mlip_train_features); returns an MLIP_* code.
int train_features(mlip_ctx* c, double* X, double* P, double* FF, double* SF);
void train_free(mlip_ctx* c);

}  // namespace mlip
