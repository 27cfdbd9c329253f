/*
 * C ABI of the MI355X-native gtinv energy/force engine.
 *
 * Drop-in boundary for the LAMMPS pair style `mlip_gtinv`
 * (src/USER-MLIP/pair_mlip_gtinv.h:17, pair_mlip_gtinv.cpp).  Every entry
 * point returns MLIP_OK or an error code and never calls exit(); the message
 * of the last failure is available from mlip_last_error().  A context is used
 * by one host thread at a time (like a Pair object) and owns one HIP stream.
 */
#ifndef MLIP_HIP_H
#define MLIP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlip_ctx mlip_ctx;
typedef struct mlip_nlist mlip_nlist;

enum {
    MLIP_OK = 0,
    MLIP_ERR_ARG = 1,     /* bad argument (e.g. unknown element, newton off, n_type > 2) */
    MLIP_ERR_FILE = 2,    /* potential file missing / malformed */
    MLIP_ERR_MODEL = 3,   /* model outside what the HIP path evaluates */
    MLIP_ERR_DEVICE = 4,  /* HIP failure (no device, allocation, launch, execution) */
    MLIP_ERR_STATE = 5,   /* call order (e.g. resident compute before resident set) */
    MLIP_ERR_FULL = 6     /* request pool: no room for the request now (retry after completions) */
};

enum { MLIP_FP64 = 0, MLIP_MIXED = 1 };

/* Replaces PairMLIPGtinv::coeff (pair_mlip_gtinv.cpp:555-593) together with
 * DataMLIPBase::initialize (mlip_model.h:35-120): `pair_coeff * * potfile
 * E1 E2 ...`.  elem_of_type[t] names the element of LAMMPS type t+1; NULL
 * entries leave a type unmapped (atoms of that type must not appear).
 * Errors that the reference reports with error->all / exit(8) / exit(1)
 * come back as MLIP_ERR_* with the message copied into errbuf. */
int mlip_create(const char* potfile, int ntypes, const char* const* elem_of_type, mlip_ctx** out,
                char* errbuf, size_t errlen);

void mlip_destroy(mlip_ctx* ctx);

/* PairMLIPGtinv::init_one (pair_mlip_gtinv.cpp:599-603): cutoff radius. */
double mlip_cutoff(const mlip_ctx* ctx);

/* Masses per LAMMPS type from the potential file (pair_mlip_gtinv.cpp:584-588). */
int mlip_masses(const mlip_ctx* ctx, double* mass_by_type);

/* PairMLIPGtinv::settings (pair_mlip_gtinv.cpp:545-549) takes no argument and
 * forces newton on; the HIP path has no counterpart beyond this check. */
int mlip_settings(int narg);

/* FP64 (default) or mixed precision (FP32 per-neighbor work, FP64 invariants). */
int mlip_set_precision(mlip_ctx* ctx, int mode);

const char* mlip_last_error(const mlip_ctx* ctx);

/* Per-atom energy and virial of mlip_compute / mlip_compute_flat:
 * MLIP_ATOM_PAIR_SPLIT (default) as the reference CPU style fills eatom and
 * vatom through ev_tally_xyz (pair_mlip_gtinv.cpp:153,240-253;
 * pair_mlip_pair.cpp:226): each pair's energy (every term of dE/da weighted
 * by 1 / its degree) and its virial del (x) f go half to each of its two
 * positions, ghosts included (their sums are LAMMPS reverse communication's);
 * MLIP_ATOM_SITE: site energy E_i of each centre and the virial of the pairs
 * it owns (no ghost entries).  Both sum to eng_vdwl and virial6. */
enum { MLIP_ATOM_PAIR_SPLIT = 0, MLIP_ATOM_SITE = 1 };
int mlip_set_atom_mode(mlip_ctx* ctx, int mode);

/* Replaces PairMLIPGtinv::compute(eflag, vflag) (pair_mlip_gtinv.cpp:63-125).
 * Reads x[nall][3] (nall = nlocal + nghost), 1-based type[nall], tag[nall]
 * and the LAMMPS neighbor list (inum, ilist, numneigh, firstneigh).
 * neigh_full = 0: half list with newton on, as the CPU style requests;
 * neigh_full = 1: full list, as the Kokkos style requests.
 * Accumulates into f[nall][3] (ghost entries are left for LAMMPS reverse
 * communication); adds to *eng_vdwl and virial6 (xx, yy, zz, xy, xz, yz).
 * eatom[nall] and vatom6[nall][6] (nullable) are accumulated as
 * mlip_set_atom_mode selects. */
int mlip_compute(mlip_ctx* ctx, int nlocal, int nghost, const double* x, const int* type,
                 const int64_t* tag, int inum, const int* ilist, const int* numneigh,
                 const int* const* firstneigh, int neigh_full, int eflag, int vflag, double* f,
                 double* eng_vdwl, double* virial6, double* eatom, double* vatom6);

/* Same contract with firstneigh flattened: the list of ilist[ii] is
 * neigh_idx[neigh_off[ii] .. neigh_off[ii] + numneigh[ilist[ii]]). */
int mlip_compute_flat(mlip_ctx* ctx, int nlocal, int nghost, const double* x, const int* type,
                      const int64_t* tag, int inum, const int* ilist, const int* numneigh,
                      const int* neigh_off, const int* neigh_idx, int neigh_full, int eflag, int vflag,
                      double* f, double* eng_vdwl, double* virial6, double* eatom, double* vatom6);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as mlip_compute_flat, bit
 * for bit, but every device buffer of the call (positions, lists, a_nlm,
 * adjoints, features, pair forces, reductions) is carved from one static
 * device array of 128 MiB per GPU, held until the process ends; the call
 * itself allocates nothing.  Structures needing more return MLIP_ERR_ARG
 * ("scratch" in the message): about 1,400 atoms of the order-4 benchmark
 * model at 137 neighbors per atom (95 KB per atom).
 * The array is shared by every context and call on a GPU and keeps the last
 * call's contents: calls that overlap on one GPU (two host threads, two
 * contexts) race. */
int mlip_compute_flat_noalloc(mlip_ctx* ctx, int nlocal, int nghost, const double* x, const int* type,
                              const int64_t* tag, int inum, const int* ilist, const int* numneigh,
                              const int* neigh_off, const int* neigh_idx, int neigh_full, int eflag,
                              int vflag, double* f, double* eng_vdwl, double* virial6, double* eatom,
                              double* vatom6);

/* ---- Device-resident evaluation (no counterpart in the reference; the
 * Kokkos style re-copies its inputs every step, pair_mlip_gtinv_kokkos_impl.h
 * :411-634).  Positions, types and a full neighbor list of owned atoms
 * with per-pair image shifts are uploaded once; mlip_resident_compute then
 * only enqueues device work on the context's stream. */
int mlip_resident_set(mlip_ctx* ctx, int natoms, const double* x, const int* type, int64_t npairs,
                      const int64_t* nbr_off, const int* nbr_idx, const double* nbr_shift);
int mlip_resident_update_positions(mlip_ctx* ctx, const double* x);
int mlip_resident_compute(mlip_ctx* ctx);
int mlip_resident_sync(mlip_ctx* ctx);
int mlip_resident_get(mlip_ctx* ctx, double* f, double* energy, double* virial6, double* eatom);

/* Resident structure from a periodic cell (cell9 = rows a, b, c): the full
 * neighbor list (0 < r < cutoff, self-images, image shifts), the force
 * gather lists and the slot layout are built on the device.  The list is
 * the one mlip_nlist_build gives, entry for entry and bit for bit, so the
 * evaluation equals mlip_resident_set with that list.  Replaces the
 * neighbor build LAMMPS runs before PairMLIPGtinv::compute (full list,
 * newton on, as requested at pair_mlip_gtinv_kokkos_impl.h:229-270; binning
 * as src/USER-MLIP/mlipkk_cell.cpp:218-307). */
int mlip_resident_set_cell(mlip_ctx* ctx, int natoms, const double* cell9, const double* x, const int* type);
/* Number of directed pairs of the resident structure or batch (-1 if none). */
int64_t mlip_resident_npairs(const mlip_ctx* ctx);
/* Copies the resident lists to host: off[natoms+1], idx[npairs],
 * shift[3 npairs] (del = x_i - x_j + shift), and the force gather lists
 * goff[natoms+1], gent[2 npairs] (for position k, entries -(p+1) for pairs
 * centred on k and p+1 for pairs with neighbor k, in increasing p).  Any
 * pointer may be NULL. */
int mlip_resident_lists(mlip_ctx* ctx, int64_t* off, int* idx, double* shift, int* goff, int* gent);

/* ---- Batched evaluation (no reference counterpart: thousands of small
 * structures per evaluation).  nstruct independent periodic structures: natoms[k] atoms
 * with cell cells[9k..9k+8] (rows a, b, c), positions x and 1-based types
 * concatenated in structure order.  Each structure gets its own periodic
 * neighbor list; one evaluation covers the whole batch and reduces energy
 * and virial per structure in a fixed order.  mlip_resident_compute /
 * mlip_resident_sync evaluate it; mlip_batch_get returns energies[nstruct],
 * virials[6 nstruct] and forces[3 natoms_total]. */
int mlip_batch_set(mlip_ctx* ctx, int nstruct, const int* natoms, const double* cells, const double* x,
                   const int* type);
int mlip_batch_get(mlip_ctx* ctx, double* energies, double* virials, double* f);
/* Training outputs (the design matrices of a fit) of the resident structure
 * or batch, gtinv descriptor: replaces MLIPModel::prepare_features and its
 * getters (src/USER-MLIP/mlipkk.cpp:461-470, 1322-1381; math
 * docs/formulation.md:282-349).  Per atom (batch order):
 * X[natoms][n_des] structural features, P[natoms][n_coeff] polynomial
 * features, FF[natoms][n_coeff][3] force features (FF[k][p] = -dX_p/dr_k
 * with X_p = sum_i P[i][p]); per structure SF[nstruct][n_coeff][6] stress
 * features in the reference's order xx, yy, zz, yz, zx, xy.  With the
 * regression weights w (coefficient / scale): sum_p w_p X_p is the energy,
 * sum_p w_p FF[k][p] the force on k and sum_p w_p SF[p] the virial.  n_des
 * and n_coeff are entries 3 and 4 of mlip_model_info.  FP64 whatever the
 * precision mode; any output may be NULL. */
int mlip_train_features(mlip_ctx* ctx, double* X, double* P, double* FF, double* SF);
/* Where mlip_batch_set builds the neighbor lists: 1 on the device (the
 * default; lists as mlip_resident_set_cell builds them), 0 on host threads
 * with the mlip_nlist_build builder.  Both give identical lists. */
int mlip_set_neighbor_builder(mlip_ctx* ctx, int device);

/* Synthetic batched/served structures: sizes log-uniform in [nmin, nmax],
 * random triclinic cells at `density` (atoms/A^3), uniform positions with an
 * rmin rejection (periodic images included), two species with a random
 * fraction.  Deterministic in seed.  Call with x = NULL to get natoms[] only,
 * then with buffers of sum(natoms) atoms. */
int mlip_random_structures(int nstruct, int nmin, int nmax, double density, double rmin, uint64_t seed,
                           int* natoms, double* cells, double* x, int* type, int64_t cap_atoms);

/* Capture one evaluation of the resident structure or batch as a HIP graph
 * and replay it on later mlip_resident_compute calls (on = 1). */
int mlip_graph_mode(mlip_ctx* ctx, int on);

/* Times each launch shape of the force kernel (64/128/192/256 threads per
 * workgroup) `reps` times on the current resident structure or batch and
 * keeps the fastest; *chosen receives it.  Also times the two a_nlm kernels
 * (where the model fits the matrix-core one) and keeps the faster. */
int mlip_autotune(mlip_ctx* ctx, int reps, int* chosen);

/* a_nlm kernel: 0 = thread per lm column with FP64 register accumulators,
 * 1 = chunk products on the matrix cores (FP64 MFMA; the default, used when
 * the model has l_max <= 12, two species and at most 32 radial functions,
 * else kernel 0 runs).  Both give the same sums up to rounding order. */
int mlip_set_anlm_kernel(mlip_ctx* ctx, int kind);
int mlip_get_anlm_kernel(const mlip_ctx* ctx);

/* Order-3/4 workgroup plan: regroups the visits over the 12 waves of the
 * order-3/4 kernel with each of a few work models (terms per scheduling
 * region and per visit) and each of three limits on the angular momentum
 * up to which the two species' rows are paired, times the adjoint kernel
 * `reps` times on the current resident structure or batch with each, and
 * keeps the fastest; plan3 receives its two constants and the pairing limit
 * (-1: no paired rows).  Results do not depend on the plan beyond the
 * summation order of split rows and of the paired rows' products. */
int mlip_autotune_q34(mlip_ctx* ctx, int reps, int* plan3);

/* Sets the order-3/4 workgroup plan directly (the three values
 * mlip_autotune_q34 returns), e.g. to reuse a plan timed before. */
int mlip_set_q34_plan(mlip_ctx* ctx, int region_cost, int visit_cost, int pair_lmax);

/* Sets the force kernel's launch shape directly (64, 128, 192 or 256
 * threads per workgroup; 192 by default).  Results do not depend on it
 * beyond the summation order of the per-atom virial. */
int mlip_set_force_threads(mlip_ctx* ctx, int threads);

/* Per-kernel device time, measured with HIP events on the context's own
 * stream around every launch while enabled.  mlip_profile_enable resets the
 * totals; mlip_profile_read returns, per kernel name, the summed
 * milliseconds and the number of launches. */
int mlip_profile_enable(mlip_ctx* ctx, int on);
int mlip_profile_read(mlip_ctx* ctx, const char** names, double* ms, int* count, int cap);

/* Algorithmic FP64 operation counts of one evaluation of the current
 * structure, per kernel: a_nlm, order 1/2, order 3/4 features, order 3/4
 * adjoint, polynomial, forces, total (see engine.hip for the rule). */
int mlip_flops(const mlip_ctx* ctx, double* out7);
/* The same seven entries by the reference-formulation rule: per half pair
 * 40 + 12 n_fn + 38 n_lmh + 8 n_fn n_lmh (a_nlm) and 44 n_fn n_lmh (force);
 * per atom n_fn sum T (6(q-1)+3) (invariants), n_fn sum T q (6(q-1)+4)
 * (adjoint), 2 sum |p| (polynomial) over the itcs / terms the centre's
 * species sees (T = CG terms of the irrep, q its order).  This is the
 * algorithmic count that roofline figures are quoted against; mlip_flops
 * counts what this implementation's kernels execute. */
int mlip_flops_reference(const mlip_ctx* ctx, double* out7);

/* Launches a plain device copy of nbytes `reps` times (byte-counter calibration). */
int mlip_calibration_copy(mlip_ctx* ctx, int64_t nbytes, int reps);

/* Diagnostics for parity tests: copy device intermediates of the last
 * evaluation to host.  what = 0: a_nlm [slot order], 1: polynomial features
 * (per slot a compact row of the features its species sees, in the engine's
 * column order), 2: per-centre energies, 3: the adjoint G (layout of a_nlm),
 * 4: the quadratic-form GEMM's output (layout of 1): the products D W, or,
 * for a model without cubic terms, whose GEMM epilogue forms the adjoint
 * weights, H = lin + D W; 5: the per-species quadratic forms, 6: eight
 * numbers (row pitch; per species the offset and dimension of its form and
 * its slot blocks; 1 if the epilogue forms H), 7: the linear weights per
 * species column ([2][row pitch]); returns the number of doubles written. */
int64_t mlip_debug_fetch(mlip_ctx* ctx, int what, double* out, int64_t cap);

/* Model description: n_type, n_fn, n_gtinv, n_des, n_coeff, maxl, lmax3,
 * n_poly_features, n_irreps, gtinv_order, model_type, maxp. */
int mlip_model_info(const mlip_ctx* ctx, int* info12);

/* Periodic full neighbor list of a single cell for the resident path (the
 * step before the path; LAMMPS builds it for the pair style).  cell9 holds
 * the rows a, b, c; pairs with 0 < r < rc including self-images, neighbor
 * indices into the owned atoms and shifts with del = x_i - x_j + shift.
 * Binning follows src/USER-MLIP/mlipkk_cell.cpp:218-307. */
int mlip_nlist_build(const double* cell9, int natoms, const double* x, double rc, mlip_nlist** out,
                     char* errbuf, size_t errlen);
int64_t mlip_nlist_npairs(const mlip_nlist* nl);
int mlip_nlist_copy(const mlip_nlist* nl, int64_t* off, int* idx, double* shift);
void mlip_nlist_free(mlip_nlist* nl);

/* Device-free model inspection for tools and tests (no reference
 * counterpart): which = 0 model info (as mlip_model_info), 1 feature
 * families (9 ints each: irrep, order, tc[6], type mask), 2 quadratic
 * products (2 feature ids each), 3 cubic products (3 each), 4 the number
 * of coefficients the model defines (the file's own count is not checked).  Returns the
 * number of ints written or a negative value with errbuf set. */
int mlip_model_inspect(const char* potfile, int which, int* out, int64_t cap, char* errbuf, size_t errlen);


/* ---- Request pool: continuous batching (serving) --------------------------
 * No counterpart in the reference (a pair style evaluates one LAMMPS
 * system); these entry points keep the per-structure semantics of
 * mlip_compute for many independent structures resident on the device.
 *
 * mlip_pool_init reserves device capacity for max_atoms atoms, max_pairs
 * directed pairs (neighbor list with skin) and max_requests requests;
 * mlip_pool_bytes gives the HBM that capacity takes, so a caller can size the
 * pool from a memory budget.  mlip_pool_admit copies one structure (cell rows
 * a, b, c; 1-based types; optional velocities in A/ps) into the pool with its
 * neighbor list built at cutoff + skin: nsteps = 0 asks for one
 * energy/force/virial evaluation, nsteps > 0 for that many velocity-Verlet
 * steps of dt ps (metal units, masses from the potential file).  It returns
 * MLIP_ERR_FULL when the capacity left cannot hold the request.
 * mlip_pool_tick enqueues one evaluation of every resident request as one
 * batch (and the NVE half steps) and returns without waiting; admissions may
 * continue while it runs and join the next tick.  mlip_pool_sync waits for
 * the tick and retires requests that finished; mlip_pool_done lists them and
 * mlip_pool_result copies one out and forgets it.  Per evaluation the pool
 * records 13 numbers: energy, virial[6] (xx, yy, zz, xy, xz, yz, eV) and the
 * kinetic-energy tensor sum m v_a v_b [6] (eV).  A request's results do not
 * depend on the other requests in its batches. */
int64_t mlip_pool_bytes(const mlip_ctx* ctx, int64_t max_atoms, int64_t max_pairs, int max_requests);
int mlip_pool_init(mlip_ctx* ctx, int64_t max_atoms, int64_t max_pairs, int max_requests);
int mlip_pool_admit(mlip_ctx* ctx, int natoms, const double* cell, const double* x, const int* type, const double* v,
                    int nsteps, double dt, double skin, int* handle);
/* Admits requests 0..count-1 in order (FIFO) until one does not fit;
 * structures concatenated (natoms[k] atoms each, cells 9 per request, v
 * NULL or concatenated).  The neighbor lists of all of them are built in
 * one batch on the device, on a stream of the pool's own that runs beside a
 * tick in flight (nthreads is accepted for compatibility and not used).
 * Returns the number admitted (negative MLIP_ERR_* on error). */
int mlip_pool_admit_many(mlip_ctx* ctx, int count, const int* natoms, const double* cells, const double* x,
                         const int* type, const double* v, const int* nsteps, double dt, const double* skin,
                         int nthreads, int* handles);
int mlip_pool_tick(mlip_ctx* ctx, int* nbatch);
int mlip_pool_sync(mlip_ctx* ctx, int* ncompleted);
/* returns the number of handles written (negative MLIP_ERR_* on error) */
int mlip_pool_done(mlip_ctx* ctx, int* handles, int cap);
/* energy/virial of the last evaluation; f, x, v [natoms*3] at the end
 * (nullable); thermo [13 per evaluation] (nullable, thermo_cap doubles).
 * Returns the number of evaluations (negative MLIP_ERR_* on error). */
int mlip_pool_result(mlip_ctx* ctx, int handle, double* energy, double* virial, double* f, double* x, double* v,
                     double* thermo, int thermo_cap);
/* used atoms, used pairs, atom capacity, pair capacity, requests held, requests running */
int mlip_pool_stats(const mlip_ctx* ctx, int64_t* out6);
/* Neighbor-list rebuilds of NVE requests so far (-1 without a pool).  An NVE
 * request needs skin > 0 (MLIP_ERR_ARG otherwise); its list is rebuilt from
 * the current positions whenever an atom would otherwise move more than
 * skin / 2 from where the list was built.  A request whose rebuilt list no
 * longer fits the pair capacity ends with an error (mlip_pool_result
 * returns -MLIP_ERR_FULL with the message in mlip_last_error). */
int64_t mlip_pool_rebuilds(const mlip_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* MLIP_HIP_H */
