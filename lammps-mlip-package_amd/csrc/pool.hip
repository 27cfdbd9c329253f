// Device-resident request pool: continuous batching of independent
// structures (single energy/force/stress evaluations and short NVE runs).
//
// Every admitted request owns an atom range and a pair range of the pool's
// device arrays (positions, velocities, element, per-centre pair ranges with
// image shifts, force-gather lists), filled once at admission from a native
// neighbor list built with a skin.  A tick evaluates all resident requests
// as one batch: the slot blocks of the pipeline mix atoms of different
// requests (same species per block), pairs and gather entries stay in each
// request's own ranges, and per-request energy, virial and kinetic-energy
// tensor come from fixed-order reductions over the request's atoms, so a
// request's numbers do not depend on what else is in the batch.  NVE
// requests advance one velocity-Verlet step per tick (LAMMPS fix nve:
// half kick and drift before the force evaluation, half kick after) in
// metal units.  Admission is by capacity: a request that does not fit the
// pool's atom or pair budget is refused with MLIP_ERR_FULL and can be
// offered again after completions free space.
//
// NVE requests need a skin: their list holds pairs within cutoff + skin of
// the positions it was built from (kept on the device as the request's
// reference positions).  At the end of every tick the device computes, per
// continuing request, the largest distance any atom will have moved from
// those reference positions after the next step's drift (the drift is known
// exactly from v and F at that point); when it exceeds skin / 2 the host
// rebuilds that request's list from the current positions before the next
// tick, so no evaluation ever runs on a list that can miss a pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <atomic>
#include <map>
#include <memory>
#include <thread>
#include <string>
#include <vector>

#include "engine_internal.h"
#include "neighbor.h"

namespace mlip {

namespace {

// LAMMPS metal units (force.cpp / update.cpp conventions)
constexpr double kMvv2e = 1.0364269e-4;        // g/mol A^2/ps^2 -> eV
constexpr double kFtm2v = 1.0 / 1.0364269e-4;  // eV/A / (g/mol) -> A/ps^2

struct NveDesc {
    int32_t a0, n;
    double dtf, dtv;  // 0.5 dt ftm2v, dt
};

__global__ void k_nve_first(const NveDesc* d, const int32_t* elem, const double* mass, const double* f, double* v,
                            double* x) {
    const NveDesc q = d[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    const int a = q.a0 + i;
    const double dtfm = q.dtf / mass[elem[a]];
    for (int c = 0; c < 3; ++c) {
        v[3 * a + c] += dtfm * f[3 * a + c];
        x[3 * a + c] += q.dtv * v[3 * a + c];
    }
}

__global__ void k_nve_final(const NveDesc* d, const int32_t* elem, const double* mass, const double* f, double* v) {
    const NveDesc q = d[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    const int a = q.a0 + i;
    const double dtfm = q.dtf / mass[elem[a]];
    for (int c = 0; c < 3; ++c) v[3 * a + c] += dtfm * f[3 * a + c];
}

// kinetic-energy tensor m v_a v_b (xx, yy, zz, xy, xz, yz) of each request,
// one workgroup per request, fixed-order tree
__global__ __launch_bounds__(256) void k_ke(const int32_t* sbeg, const int32_t* send, const int32_t* elem,
                                            const double* mass, const double* v, double* out) {
    __shared__ double sh[256];
    const int r = blockIdx.x;
    const int a0 = sbeg[r], a1 = send[r];
    for (int c = 0; c < 6; ++c) {
        const int p = c < 3 ? c : (c == 5 ? 1 : 0);
        const int q = c < 3 ? c : (c == 3 ? 1 : 2);
        double s = 0.0;
        for (int a = a0 + threadIdx.x; a < a1; a += 256) s += mass[elem[a]] * v[3 * a + p] * v[3 * a + q];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[6 * r + c] = kMvv2e * sh[0];
        __syncthreads();
    }
}

// Largest displacement from the list's reference positions after the next
// drift x + dt (v + dtf F / m), per batch request (n = 0: no next step)
__global__ __launch_bounds__(256) void k_disp_next(const NveDesc* d, const int32_t* elem, const double* mass,
                                                   const double* f, const double* v, const double* x,
                                                   const double* xref, double* out) {
    __shared__ double sh[256];
    const NveDesc q = d[blockIdx.x];
    double m2 = 0.0;
    for (int i = threadIdx.x; i < q.n; i += 256) {
        const int a = q.a0 + i;
        const double dtfm = q.dtf / mass[elem[a]];
        double d2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double xn = x[3 * a + c] + q.dtv * (v[3 * a + c] + dtfm * f[3 * a + c]);
            const double dd = xn - xref[3 * a + c];
            d2 += dd * dd;
        }
        m2 = fmax(m2, d2);
    }
    sh[threadIdx.x] = m2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sqrt(sh[0]);
}

__global__ void k_scatter_slots(const int32_t* atom_of_slot, int nslot, int32_t* slot_of_atom) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const int a = atom_of_slot[s];
    if (a >= 0) slot_of_atom[a] = s;
}

// Packed host->device staging: one copy of a pinned arena, then this kernel
// moves every piece to its destination (32-bit words; destinations are
// 4-byte aligned, pieces 16-byte aligned in the arena).
struct ScatterDesc {
    uint64_t src;  // byte offset in the staged arena
    uint64_t dst;  // device address
    uint64_t nwords;
};

__global__ void k_scatter(const ScatterDesc* d, const unsigned char* stage) {
    const ScatterDesc q = d[blockIdx.x];
    const unsigned* src = reinterpret_cast<const unsigned*>(stage + q.src);
    unsigned* dst = reinterpret_cast<unsigned*>(q.dst);
    const uint64_t stride = static_cast<uint64_t>(gridDim.y) * blockDim.x;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.y) * blockDim.x + threadIdx.x; w < q.nwords; w += stride)
        dst[w] = src[w];
}

struct HostArena {
    unsigned char* h = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};

// First-fit allocator of index ranges with coalescing of freed ranges.
class RangeAlloc {
  public:
    void init(int64_t cap) {
        cap_ = cap;
        used_ = 0;
        free_.clear();
        if (cap > 0) free_[0] = cap;
    }
    int64_t take(int64_t n) {
        if (n <= 0) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const int64_t off = it->first, rem = it->second - n;
            free_.erase(it);
            if (rem > 0) free_[off + n] = rem;
            used_ += n;
            return off;
        }
        return -1;
    }
    void give(int64_t off, int64_t n) {
        if (n <= 0) return;
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t used() const { return used_; }
    int64_t cap() const { return cap_; }
    // end of the highest allocated range
    int64_t high() const {
        if (free_.empty()) return cap_;
        const auto& last = *free_.rbegin();
        return last.first + last.second == cap_ ? last.first : cap_;
    }

  private:
    int64_t cap_ = 0, used_ = 0;
    std::map<int64_t, int64_t> free_;
};

struct Request {
    int id = 0;
    int natoms = 0;
    int64_t a0 = 0, p0 = 0, np = 0;
    int nsteps = 0, step = 0;
    double dt = 0.0, skin = 0.0;
    double cell[9] = {};
    int rebuilds = 0;
    std::vector<int32_t> elem;
    bool done = false;
    std::string error;  // set when the request failed (done with no result)
    // per evaluation: energy, virial[6], kinetic-energy tensor[6]
    std::vector<double> thermo;
    std::vector<double> f, x, v;
};

constexpr int kThermo = 13;

}  // namespace

struct Pool {
    int64_t cap_atoms = 0, cap_pairs = 0, cap_slots = 0;
    int cap_req = 0;
    RangeAlloc atoms, pairs;
    DBuf<double> x, v, f, pshift, A, G, D, Hq, H, Ep, Eq12, Eq34, gpair, vown, e_slot, red, out, sout, ke, mass;
    DBuf<int32_t> elem, pbeg, pend, pnbr, gbeg, gend, gent, atom_of_slot, slot_of_atom, block_type, sbeg, send;
    DBuf<NveDesc> first, final_, disp;
    DBuf<double> xref, dnext;  // list reference positions; next-step displacement per batch request
    int64_t rebuilds = 0;      // neighbor-list rebuilds of NVE requests so far
    std::map<int, Request> req;
    int next_id = 1;
    std::vector<int> done_ids;
    // the tick in flight: requests in batch order and host staging
    bool in_flight = false;
    std::vector<int> batch;
    double* h_out = nullptr;  // pinned: 7 per request, 6 kinetic per request, 1 next-step displacement
    // packed uploads: two pinned arenas used in turn, one device copy
    HostArena arena[2];
    int cur = 0;
    size_t fill = 0;
    std::vector<ScatterDesc> descs;
    DBuf<unsigned char> dstage;
    DBuf<ScatterDesc> ddesc;
};

void pool_free(mlip_ctx* c) {
    Pool* p = c->pool;
    if (!p) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DBuf<double>* dbs[] = {&p->x,  &p->v,      &p->f,    &p->pshift, &p->A,   &p->G,   &p->D,
                           &p->Hq, &p->H,      &p->Ep,   &p->Eq12,   &p->Eq34, &p->gpair, &p->vown,
                           &p->e_slot, &p->red, &p->out, &p->sout,   &p->ke,  &p->mass};
    for (auto* b : dbs) b->release();
    DBuf<int32_t>* ibs[] = {&p->elem, &p->pbeg, &p->pend, &p->pnbr, &p->gbeg, &p->gend, &p->gent,
                            &p->atom_of_slot, &p->slot_of_atom, &p->block_type, &p->sbeg, &p->send};
    for (auto* b : ibs) b->release();
    p->first.release();
    p->final_.release();
    p->disp.release();
    p->xref.release();
    p->dnext.release();
    p->dstage.release();
    p->ddesc.release();
    for (auto& a : p->arena) {
        if (a.h) (void)hipHostFree(a.h);
        if (a.ev) (void)hipEventDestroy(a.ev);
    }
    if (p->h_out) (void)hipHostFree(p->h_out);
    delete p;
    c->pool = nullptr;
}

namespace {

// start a packed upload (waits until the arena's previous copy is done)
bool stage_begin(mlip_ctx* c, Pool* p) {
    HostArena& a = p->arena[p->cur];
    if (a.pending && !hip_ok(c, hipEventSynchronize(a.ev), "staging arena")) return false;
    a.pending = false;
    p->fill = 0;
    p->descs.clear();
    return true;
}

bool arena_reserve(mlip_ctx* c, HostArena& a, size_t need, size_t keep) {
    if (need <= a.cap) return true;
    size_t cap = std::max<size_t>(need, 2 * a.cap);
    cap = std::max<size_t>(cap, 1 << 22);
    unsigned char* h = nullptr;
    if (!hip_ok(c, hipHostMalloc(reinterpret_cast<void**>(&h), cap), "pinned staging")) return false;
    if (a.h) {
        std::memcpy(h, a.h, keep);
        (void)hipHostFree(a.h);
    }
    a.h = h;
    a.cap = cap;
    return true;
}

template <typename T>
bool stage_upload(mlip_ctx* c, Pool* p, T* dst, const T* src, size_t n, const char*) {
    static_assert(sizeof(T) % 4 == 0, "staged pieces move as 32-bit words");
    const size_t nbytes = n * sizeof(T);
    if (nbytes == 0) return true;
    HostArena& a = p->arena[p->cur];
    const size_t off = (p->fill + 15) & ~static_cast<size_t>(15);
    if (!arena_reserve(c, a, off + nbytes, p->fill)) return false;
    std::memcpy(a.h + off, src, nbytes);
    p->fill = off + nbytes;
    p->descs.push_back({off, reinterpret_cast<uint64_t>(dst), (nbytes + 3) / 4});
    return true;
}

// one device copy of the arena, then the scatter kernel
bool stage_flush(mlip_ctx* c, Pool* p) {
    if (p->descs.empty()) return true;
    HostArena& a = p->arena[p->cur];
    if (!a.ev && !hip_ok(c, hipEventCreateWithFlags(&a.ev, hipEventDisableTiming), "staging event")) return false;
    const size_t doff = (p->fill + 15) & ~static_cast<size_t>(15);
    const size_t dbytes = p->descs.size() * sizeof(ScatterDesc);
    if (!arena_reserve(c, a, doff + dbytes, p->fill)) return false;
    std::memcpy(a.h + doff, p->descs.data(), dbytes);
    const size_t total = doff + dbytes;
    if (total > p->dstage.cap) {
        // growing the device stage frees the old one; wait for its readers
        if (!hip_ok(c, hipStreamSynchronize(c->stream), "staging grow")) return false;
        if (!p->dstage.ensure(std::max<size_t>(total, 2 * p->dstage.cap))) {
            c->err = "device allocation failed for staging";
            return false;
        }
    }
    if (!hip_ok(c, hipMemcpyAsync(p->dstage.p, a.h, total, hipMemcpyHostToDevice, c->stream), "staged upload"))
        return false;
    hipLaunchKernelGGL(k_scatter, dim3(static_cast<unsigned>(p->descs.size()), 8), dim3(256), 0, c->stream,
                       reinterpret_cast<const ScatterDesc*>(p->dstage.p + doff), p->dstage.p);
    if (!hip_ok(c, hipGetLastError(), "scatter launch") || !hip_ok(c, hipEventRecord(a.ev, c->stream), "staging event"))
        return false;
    a.pending = true;
    p->cur ^= 1;
    p->descs.clear();
    p->fill = 0;
    return true;
}

// point the pipeline's system view at the pool arrays for one tick
void bind_pool(mlip_ctx* c, Pool* p, int nblocks) {
    DevSystem& s = c->ds;
    std::memset(&s, 0, sizeof s);
    s.nall = static_cast<int>(p->atoms.high());
    s.nown = s.nall;
    s.nblocks = nblocks;
    s.npairs = p->pairs.high();
    s.x = p->x.p;
    s.elem = p->elem.p;
    s.poff = p->pbeg.p;
    s.pend = p->pend.p;
    s.pnbr = p->pnbr.p;
    s.pcen = nullptr;
    s.pshift = p->pshift.p;
    s.atom_of_slot = p->atom_of_slot.p;
    s.slot_of_atom = p->slot_of_atom.p;
    s.block_type = p->block_type.p;
    s.A = p->A.p;
    s.G = p->G.p;
    s.D = p->D.p;
    s.Hq = p->Hq.p;
    s.H = p->H.p;
    s.Ep = p->Ep.p;
    s.Eq12 = p->Eq12.p;
    s.Eq34 = p->Eq34.p;
    s.gpair = p->gpair.p;
    s.vown = p->vown.p;
    s.goff = p->gbeg.p;
    s.gend = p->gend.p;
    s.gent = p->gent.p;
    s.f = p->f.p;
    s.e_slot = p->e_slot.p;
    s.red = p->red.p;
    s.out = p->out.p;
    c->npairs = s.npairs;
}

}  // namespace

}  // namespace mlip

using namespace mlip;

extern "C" {

int mlip_pool_init(mlip_ctx* c, int64_t max_atoms, int64_t max_pairs, int max_requests) {
    if (!c || max_atoms < 1 || max_pairs < 1 || max_requests < 1 || max_atoms > INT32_MAX / 4 ||
        max_pairs > INT32_MAX / 4)
        return MLIP_ERR_ARG;
    pool_free(c);
    drop_graph(c);
    auto* p = new Pool();
    c->pool = p;
    const Model& m = c->model;
    p->cap_atoms = max_atoms;
    p->cap_pairs = max_pairs;
    p->cap_req = max_requests;
    // every species pads its last slot block
    p->cap_slots = (max_atoms + static_cast<int64_t>(m.n_type) * kBlock + kBlock - 1) / kBlock * kBlock;
    p->atoms.init(max_atoms);
    p->pairs.init(max_pairs);
    const size_t na = static_cast<size_t>(max_atoms), npr = static_cast<size_t>(max_pairs);
    const size_t ns = static_cast<size_t>(p->cap_slots), nr = static_cast<size_t>(max_requests);
    const size_t asz = ns * m.n_fn * m.n_type * m.n_lmh * 2, np = static_cast<size_t>(m.n_p);
    const bool ok = p->x.ensure(3 * na) && p->v.ensure(3 * na) && p->f.ensure(3 * na) && p->elem.ensure(na) &&
                    p->pbeg.ensure(na) && p->pend.ensure(na) && p->gbeg.ensure(na) && p->gend.ensure(na) &&
                    p->slot_of_atom.ensure(na) && p->vown.ensure(6 * na) && p->pnbr.ensure(npr) &&
                    p->pshift.ensure(3 * npr) && p->gpair.ensure(3 * npr) && p->gent.ensure(2 * npr) &&
                    p->A.ensure(asz) && p->G.ensure(asz) && p->D.ensure(ns * np) && p->Hq.ensure(ns * np) &&
                    p->H.ensure(ns * np) && p->Ep.ensure(ns) && p->Eq12.ensure(ns * m.n_fn) &&
                    p->Eq34.ensure(ns * m.n_fn * kWavesPerWg) && p->e_slot.ensure(ns) &&
                    p->atom_of_slot.ensure(ns) && p->block_type.ensure(ns / kBlock + 1) && p->red.ensure(7 * 128) &&
                    p->out.ensure(8) && p->sout.ensure(7 * nr) && p->ke.ensure(6 * nr) && p->sbeg.ensure(nr) &&
                    p->send.ensure(nr) && p->first.ensure(nr) && p->final_.ensure(nr) && p->disp.ensure(nr) &&
                    p->dnext.ensure(nr) && p->xref.ensure(3 * na) && p->mass.ensure(static_cast<size_t>(m.n_type));
    if (!ok || hipHostMalloc(reinterpret_cast<void**>(&p->h_out), 14 * nr * sizeof(double)) != hipSuccess) {
        c->err = "device allocation failed for the request pool";
        pool_free(c);
        return MLIP_ERR_DEVICE;
    }
    // empty ranges everywhere; a_nlm of padding lanes starts finite
    std::vector<double> mass(m.masses.begin(), m.masses.end());
    mass.resize(m.n_type, 1.0);
    if (!hip_ok(c, hipMemsetAsync(p->pbeg.p, 0, na * 4, c->stream), "pool clear") ||
        !hip_ok(c, hipMemsetAsync(p->pend.p, 0, na * 4, c->stream), "pool clear") ||
        !hip_ok(c, hipMemsetAsync(p->gbeg.p, 0, na * 4, c->stream), "pool clear") ||
        !hip_ok(c, hipMemsetAsync(p->gend.p, 0, na * 4, c->stream), "pool clear") ||
        !hip_ok(c, hipMemsetAsync(p->elem.p, 0, na * 4, c->stream), "pool clear") ||
        !hip_ok(c, hipMemsetAsync(p->x.p, 0, 3 * na * 8, c->stream), "pool clear") ||
        !hip_ok(c, hipMemsetAsync(p->A.p, 0, asz * 8, c->stream), "pool clear") ||
        !hip_ok(c, hipMemcpyAsync(p->mass.p, mass.data(), mass.size() * 8, hipMemcpyHostToDevice, c->stream),
                "pool masses") ||
        !hip_ok(c, hipStreamSynchronize(c->stream), "pool init")) {
        pool_free(c);
        return MLIP_ERR_DEVICE;
    }
    c->resident = false;
    return MLIP_OK;
}

}  // extern "C"

namespace mlip {
namespace {

// Host half of an admission that needs no pool state (runs on worker
// threads): element mapping, the neighbor list at cutoff + skin and the
// request-local force-gather lists.
struct Prepared {
    int rc = MLIP_OK;
    std::string err;
    std::vector<int32_t> elem;
    NeighborList nl;
    std::vector<int32_t> loff, goff, gent;
    int64_t a0 = -1, p0 = -1;        // ranges, once placed
    size_t off[11] = {};             // arena offsets of the eleven pieces
};

void prepare(const mlip_ctx* c, int natoms, const double* cell, const double* x, const int* type, double skin,
             Prepared& out, int inner_threads = 1) {
    out.elem.resize(natoms);
    for (int i = 0; i < natoms; ++i) {
        const int t = type[i];
        if (t < 1 || t > static_cast<int>(c->elem_of_type.size()) || c->elem_of_type[t - 1] < 0) {
            out.rc = MLIP_ERR_ARG;
            out.err = "atom type without an element mapping";
            return;
        }
        out.elem[i] = c->elem_of_type[t - 1];
    }
    std::string e;
    if (!build_full_list(cell, natoms, x, c->model.cutoff + skin, out.nl, e, inner_threads)) {
        out.rc = MLIP_ERR_ARG;
        out.err = "request structure: " + e;
        return;
    }
    out.loff.resize(natoms + 1);
    for (int i = 0; i <= natoms; ++i) out.loff[i] = static_cast<int32_t>(out.nl.off[i]);
    build_gather(natoms, out.loff, out.nl.idx, nullptr, out.goff, out.gent);
}

template <typename F>
void parallel_for(int count, int nthreads, F&& f) {
    const int nt = std::max(1, std::min(nthreads, count));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k = next++; k < count; k = next++) f(k);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// Admits prepared requests 0..count-1 in order until one does not fit:
// ranges are taken sequentially (FIFO), then every admitted request's
// pieces are written into the pinned arena by worker threads and one packed
// copy moves them.  Returns the number admitted, or -MLIP_ERR_*.
int commit_many(mlip_ctx* c, Pool* p, std::vector<Prepared>& pr, int count, const int* natoms, const int64_t* at,
                const double* cells, const double* x, const double* v, const int* nsteps, double dt,
                const double* skins, int nthreads, int* handles) {
    int n = 0;
    for (; n < count; ++n) {
        Prepared& q = pr[n];
        if (q.rc != MLIP_OK) {
            c->err = q.err;
            if (n == 0) return -q.rc;
            break;
        }
        if (static_cast<int>(p->req.size()) + n >= p->cap_req) break;
        const int64_t np = static_cast<int64_t>(q.nl.idx.size());
        q.a0 = p->atoms.take(natoms[n]);
        if (q.a0 < 0) break;
        q.p0 = p->pairs.take(std::max<int64_t>(np, 1));
        if (q.p0 < 0) {
            p->atoms.give(q.a0, natoms[n]);
            q.a0 = -1;
            break;
        }
    }
    if (n == 0) {
        c->err = "request pool: capacity exhausted";
        return 0;
    }
    // arena layout: eleven 16-byte-aligned pieces per request
    if (!stage_begin(c, p)) return -MLIP_ERR_DEVICE;
    HostArena& ar = p->arena[p->cur];
    size_t fill = 0;
    auto place = [&](size_t bytes) {
        const size_t o = (fill + 15) & ~static_cast<size_t>(15);
        fill = o + bytes;
        return o;
    };
    for (int k = 0; k < n; ++k) {
        Prepared& q = pr[k];
        const size_t na = static_cast<size_t>(natoms[k]), np = q.nl.idx.size();
        const size_t bytes[11] = {24 * na, 24 * na, 4 * na, 4 * na, 4 * na, 4 * na, 4 * na, 4 * np, 24 * np, 8 * np,
                                  24 * na};
        double* dx = p->x.p + 3 * q.a0;
        void* dst[11] = {dx,
                         p->v.p + 3 * q.a0,
                         p->elem.p + q.a0,
                         p->pbeg.p + q.a0,
                         p->pend.p + q.a0,
                         p->gbeg.p + q.a0,
                         p->gend.p + q.a0,
                         p->pnbr.p + q.p0,
                         p->pshift.p + 3 * q.p0,
                         p->gent.p + 2 * q.p0,
                         p->xref.p + 3 * q.a0};
        for (int j = 0; j < 11; ++j) {
            q.off[j] = place(bytes[j]);
            if (bytes[j]) p->descs.push_back({q.off[j], reinterpret_cast<uint64_t>(dst[j]), bytes[j] / 4});
        }
    }
    if (!arena_reserve(c, ar, fill, 0)) return -MLIP_ERR_DEVICE;
    parallel_for(n, nthreads, [&](int k) {
        Prepared& q = pr[k];
        const int na = natoms[k];
        const int64_t np = static_cast<int64_t>(q.nl.idx.size());
        unsigned char* h = ar.h;
        std::memcpy(h + q.off[0], x + 3 * at[k], 24 * static_cast<size_t>(na));
        std::memcpy(h + q.off[10], x + 3 * at[k], 24 * static_cast<size_t>(na));
        if (v)
            std::memcpy(h + q.off[1], v + 3 * at[k], 24 * static_cast<size_t>(na));
        else
            std::memset(h + q.off[1], 0, 24 * static_cast<size_t>(na));
        std::memcpy(h + q.off[2], q.elem.data(), 4 * static_cast<size_t>(na));
        int32_t* pb = reinterpret_cast<int32_t*>(h + q.off[3]);
        int32_t* pe = reinterpret_cast<int32_t*>(h + q.off[4]);
        int32_t* gb = reinterpret_cast<int32_t*>(h + q.off[5]);
        int32_t* ge = reinterpret_cast<int32_t*>(h + q.off[6]);
        const int64_t g0 = 2 * q.p0;
        for (int i = 0; i < na; ++i) {
            pb[i] = static_cast<int32_t>(q.p0 + q.loff[i]);
            pe[i] = static_cast<int32_t>(q.p0 + q.loff[i + 1]);
            gb[i] = static_cast<int32_t>(g0 + q.goff[i]);
            ge[i] = static_cast<int32_t>(g0 + q.goff[i + 1]);
        }
        int32_t* nb = reinterpret_cast<int32_t*>(h + q.off[7]);
        for (int64_t e = 0; e < np; ++e) nb[e] = static_cast<int32_t>(q.a0 + q.nl.idx[e]);
        std::memcpy(h + q.off[8], q.nl.shift.data(), 24 * static_cast<size_t>(np));
        int32_t* gt = reinterpret_cast<int32_t*>(h + q.off[9]);
        for (size_t e = 0; e < q.gent.size(); ++e) {
            const int32_t en = q.gent[e];
            gt[e] = en > 0 ? static_cast<int32_t>(en + q.p0) : static_cast<int32_t>(en - q.p0);
        }
    });
    p->fill = fill;
    if (!stage_flush(c, p)) return -MLIP_ERR_DEVICE;
    for (int k = 0; k < n; ++k) {
        Prepared& q = pr[k];
        Request r;
        r.natoms = natoms[k];
        r.nsteps = nsteps[k];
        r.dt = dt;
        r.skin = skins[k];
        std::copy(cells + 9 * k, cells + 9 * k + 9, r.cell);
        r.a0 = q.a0;
        r.p0 = q.p0;
        r.np = static_cast<int64_t>(q.nl.idx.size());
        r.elem = std::move(q.elem);
        r.id = p->next_id++;
        handles[k] = r.id;
        p->req.emplace(r.id, std::move(r));
    }
    return n;
}

int default_threads() {
    return static_cast<int>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
}

// Frees a request's device ranges (its atoms own no pairs or gather entries
// any more).
bool release_request(mlip_ctx* c, Pool* p, Request& r) {
    const size_t a0 = static_cast<size_t>(r.a0), n = static_cast<size_t>(r.natoms);
    if (!hip_ok(c, hipMemsetAsync(p->pbeg.p + a0, 0, n * 4, c->stream), "pool release") ||
        !hip_ok(c, hipMemsetAsync(p->pend.p + a0, 0, n * 4, c->stream), "pool release") ||
        !hip_ok(c, hipMemsetAsync(p->gbeg.p + a0, 0, n * 4, c->stream), "pool release") ||
        !hip_ok(c, hipMemsetAsync(p->gend.p + a0, 0, n * 4, c->stream), "pool release"))
        return false;
    p->atoms.give(r.a0, r.natoms);
    p->pairs.give(r.p0, std::max<int64_t>(r.np, 1));
    return true;
}

// New neighbor lists (cutoff + skin, from the current positions) for NVE
// requests whose atoms would pass skin / 2 in their next step; runs between
// ticks (the stream is idle).  A request whose new list does not fit the
// pair capacity fails with an error (reported by mlip_pool_result).
int rebuild_lists(mlip_ctx* c, Pool* p, const std::vector<int>& ids, int* ndone) {
    struct Job {
        Request* r;
        std::vector<double> x;
        NeighborList nl;
        std::vector<int32_t> loff, goff, gent;
        std::string err;
        bool ok = true;
    };
    std::vector<Job> jobs(ids.size());
    for (size_t k = 0; k < ids.size(); ++k) {
        Request& r = p->req[ids[k]];
        jobs[k].r = &r;
        jobs[k].x.resize(3 * static_cast<size_t>(r.natoms));
        if (!hip_ok(c, hipMemcpy(jobs[k].x.data(), p->x.p + 3 * r.a0, jobs[k].x.size() * 8, hipMemcpyDeviceToHost),
                    "pool positions for a list rebuild"))
            return MLIP_ERR_DEVICE;
    }
    const double rc = c->model.cutoff;
    parallel_for(static_cast<int>(jobs.size()), default_threads(), [&](int k) {
        Job& j = jobs[k];
        j.ok = build_full_list(j.r->cell, j.r->natoms, j.x.data(), rc + j.r->skin, j.nl, j.err, 1);
        if (!j.ok) return;
        j.loff.resize(j.r->natoms + 1);
        for (int i = 0; i <= j.r->natoms; ++i) j.loff[i] = static_cast<int32_t>(j.nl.off[i]);
        build_gather(j.r->natoms, j.loff, j.nl.idx, nullptr, j.goff, j.gent);
    });
    if (!stage_begin(c, p)) return MLIP_ERR_DEVICE;
    for (Job& j : jobs) {
        Request& r = *j.r;
        const int64_t np = j.ok ? static_cast<int64_t>(j.nl.idx.size()) : 0;
        p->pairs.give(r.p0, std::max<int64_t>(r.np, 1));
        const int64_t p0 = j.ok ? p->pairs.take(std::max<int64_t>(np, 1)) : -1;
        if (p0 < 0) {
            // no list for the request: it ends here with an error
            r.error = j.ok ? "request pool: no pair capacity for an NVE neighbor-list rebuild"
                           : "NVE neighbor-list rebuild: " + j.err;
            r.np = 0;
            r.p0 = p->pairs.take(1);
            r.done = true;
            if (!release_request(c, p, r)) return MLIP_ERR_DEVICE;
            p->done_ids.push_back(r.id);
            ++*ndone;
            continue;
        }
        r.p0 = p0;
        r.np = np;
        ++r.rebuilds;
        ++p->rebuilds;
        const int na = r.natoms;
        std::vector<int32_t> pb(na), pe(na), gb(na), ge(na), nb(static_cast<size_t>(np)), gt(j.gent.size());
        const int64_t g0 = 2 * p0;
        for (int i = 0; i < na; ++i) {
            pb[i] = static_cast<int32_t>(p0 + j.loff[i]);
            pe[i] = static_cast<int32_t>(p0 + j.loff[i + 1]);
            gb[i] = static_cast<int32_t>(g0 + j.goff[i]);
            ge[i] = static_cast<int32_t>(g0 + j.goff[i + 1]);
        }
        for (int64_t e = 0; e < np; ++e) nb[e] = static_cast<int32_t>(r.a0 + j.nl.idx[e]);
        for (size_t e = 0; e < j.gent.size(); ++e) {
            const int32_t en = j.gent[e];
            gt[e] = en > 0 ? static_cast<int32_t>(en + p0) : static_cast<int32_t>(en - p0);
        }
        const size_t a3 = 3 * static_cast<size_t>(r.a0);
        if (!stage_upload(c, p, p->pbeg.p + r.a0, pb.data(), pb.size(), "rebuild") ||
            !stage_upload(c, p, p->pend.p + r.a0, pe.data(), pe.size(), "rebuild") ||
            !stage_upload(c, p, p->gbeg.p + r.a0, gb.data(), gb.size(), "rebuild") ||
            !stage_upload(c, p, p->gend.p + r.a0, ge.data(), ge.size(), "rebuild") ||
            !stage_upload(c, p, p->pnbr.p + p0, nb.data(), nb.size(), "rebuild") ||
            !stage_upload(c, p, p->pshift.p + 3 * p0, j.nl.shift.data(), j.nl.shift.size(), "rebuild") ||
            !stage_upload(c, p, p->gent.p + 2 * p0, gt.data(), gt.size(), "rebuild") ||
            !stage_upload(c, p, p->xref.p + a3, j.x.data(), j.x.size(), "rebuild"))
            return MLIP_ERR_DEVICE;
    }
    return stage_flush(c, p) ? MLIP_OK : MLIP_ERR_DEVICE;
}

}  // namespace
}  // namespace mlip

extern "C" {

int mlip_pool_admit(mlip_ctx* c, int natoms, const double* cell, const double* x, const int* type, const double* v,
                    int nsteps, double dt, double skin, int* handle) {
    if (!c || !c->pool || natoms < 1 || !cell || !x || !type || !handle || nsteps < 0 || skin < 0.0 ||
        (nsteps > 0 && !(dt > 0.0)))
        return MLIP_ERR_ARG;
    if (nsteps > 0 && !(skin > 0.0)) {
        c->err = "NVE requests need a neighbor skin > 0 (the list is rebuilt when an atom moves skin / 2)";
        return MLIP_ERR_ARG;
    }
    std::vector<Prepared> pr(1);
    prepare(c, natoms, cell, x, type, skin, pr[0], default_threads());
    const int64_t at[2] = {0, natoms};
    const int k = commit_many(c, c->pool, pr, 1, &natoms, at, cell, x, v, &nsteps, dt, &skin, 1, handle);
    if (k < 0) return -k;
    return k == 1 ? MLIP_OK : MLIP_ERR_FULL;
}

int mlip_pool_admit_many(mlip_ctx* c, int count, const int* natoms, const double* cells, const double* x,
                         const int* type, const double* v, const int* nsteps, double dt, const double* skin,
                         int nthreads, int* handles) {
    if (!c || !c->pool || count < 0 || (count > 0 && (!natoms || !cells || !x || !type || !nsteps || !skin ||
                                                      !handles)))
        return -MLIP_ERR_ARG;
    if (count == 0) return 0;
    std::vector<int64_t> at(count + 1, 0);
    for (int k = 0; k < count; ++k) {
        if (natoms[k] < 1 || nsteps[k] < 0 || skin[k] < 0.0 || (nsteps[k] > 0 && !(dt > 0.0))) return -MLIP_ERR_ARG;
        if (nsteps[k] > 0 && !(skin[k] > 0.0)) {
            c->err = "NVE requests need a neighbor skin > 0 (the list is rebuilt when an atom moves skin / 2)";
            return -MLIP_ERR_ARG;
        }
        at[k + 1] = at[k] + natoms[k];
    }
    const int nt = nthreads > 0 ? nthreads : default_threads();
    std::vector<Prepared> pr(count);
    // few requests: threads inside each neighbor build; many: one per request
    const int inner = std::max(1, nt / count);
    parallel_for(count, std::min(nt, count), [&](int k) {
        prepare(c, natoms[k], cells + 9 * k, x + 3 * at[k], type + at[k], skin[k], pr[k], inner);
    });
    return commit_many(c, c->pool, pr, count, natoms, at.data(), cells, x, v, nsteps, dt, skin, nt, handles);
}

int mlip_pool_tick(mlip_ctx* c, int* nbatch) {
    if (!c || !c->pool) return MLIP_ERR_ARG;
    Pool* p = c->pool;
    if (p->in_flight) {
        c->err = "pool tick already in flight (call mlip_pool_sync)";
        return MLIP_ERR_STATE;
    }
    p->batch.clear();
    for (const auto& kv : p->req)
        if (!kv.second.done) p->batch.push_back(kv.first);
    if (nbatch) *nbatch = static_cast<int>(p->batch.size());
    if (p->batch.empty()) return MLIP_OK;
    const Model& m = c->model;
    const int nb = static_cast<int>(p->batch.size());
    // slot blocks by species over the batch, request ranges, NVE steps
    std::vector<int32_t> aos, btype, sbeg(nb), send(nb);
    std::vector<NveDesc> first, final_, disp(nb);
    for (int t = 0; t < m.n_type; ++t) {
        size_t k = 0;
        for (int id : p->batch) {
            const Request& r = p->req[id];
            for (int i = 0; i < r.natoms; ++i)
                if (r.elem[i] == t) {
                    if (k % kBlock == 0) btype.push_back(t);
                    aos.push_back(static_cast<int32_t>(r.a0 + i));
                    ++k;
                }
        }
        while (k % kBlock) {
            aos.push_back(-1);
            ++k;
        }
    }
    for (int b = 0; b < nb; ++b) {
        const Request& r = p->req[p->batch[b]];
        sbeg[b] = static_cast<int32_t>(r.a0);
        send[b] = static_cast<int32_t>(r.a0 + r.natoms);
        const NveDesc d{static_cast<int32_t>(r.a0), r.natoms, 0.5 * r.dt * kFtm2v, r.dt};
        if (r.nsteps > 0 && r.step > 0) {
            first.push_back(d);
            final_.push_back(d);
        }
        // a step follows this tick: check the displacement it will reach
        disp[b] = d;
        if (!(r.nsteps > 0 && r.step < r.nsteps)) disp[b].n = 0;
    }
    const int nslot = static_cast<int>(aos.size());
    if (nslot > p->cap_slots) {
        c->err = "request pool: slot capacity exceeded";
        return MLIP_ERR_STATE;
    }
    int maxn = 0;
    for (const auto& d : first) maxn = std::max(maxn, d.n);
    bool ok = stage_begin(c, p) &&
              stage_upload(c, p, p->atom_of_slot.p, aos.data(), aos.size(), "pool slots") &&
              stage_upload(c, p, p->block_type.p, btype.data(), btype.size(), "pool block types") &&
              stage_upload(c, p, p->sbeg.p, sbeg.data(), sbeg.size(), "pool request ranges") &&
              stage_upload(c, p, p->send.p, send.data(), send.size(), "pool request ranges") &&
              stage_upload(c, p, p->first.p, first.data(), first.size(), "pool nve") &&
              stage_upload(c, p, p->final_.p, final_.data(), final_.size(), "pool nve") &&
              stage_upload(c, p, p->disp.p, disp.data(), disp.size(), "pool displacement") && stage_flush(c, p);
    if (!ok) return MLIP_ERR_DEVICE;
    hipStream_t st = c->stream;
    if (!first.empty())
        hipLaunchKernelGGL(k_nve_first, dim3((maxn + 255) / 256, static_cast<unsigned>(first.size())), dim3(256), 0,
                           st, p->first.p, p->elem.p, p->mass.p, p->f.p, p->v.p, p->x.p);
    hipLaunchKernelGGL(k_scatter_slots, dim3((nslot + 255) / 256), dim3(256), 0, st, p->atom_of_slot.p, nslot,
                       p->slot_of_atom.p);
    if (!hip_ok(c, launched(), "tick launch")) return MLIP_ERR_DEVICE;
    c->nblk_type[0] = c->nblk_type[1] = 0;
    for (int32_t t : btype) ++c->nblk_type[t];
    bind_pool(c, p, static_cast<int>(btype.size()));
    c->nstruct = nb;
    c->sbeg = p->sbeg.p;
    c->send = p->send.p;
    c->sres = p->sout.p;
    c->total_reduce = false;
    // work accounting for mlip_flops
    c->owners_per_type.assign(m.n_type, 0);
    int64_t npr = 0;
    for (int id : p->batch) {
        for (int32_t e : p->req[id].elem) ++c->owners_per_type[e];
        npr += p->req[id].np;
    }
    c->npairs = npr;
    if (!run_pipeline(c)) return MLIP_ERR_DEVICE;
    if (!final_.empty())
        hipLaunchKernelGGL(k_nve_final, dim3((maxn + 255) / 256, static_cast<unsigned>(final_.size())), dim3(256), 0,
                           st, p->final_.p, p->elem.p, p->mass.p, p->f.p, p->v.p);
    if (!hip_ok(c, launched(), "nve launch")) return MLIP_ERR_DEVICE;
    hipLaunchKernelGGL(k_ke, dim3(nb), dim3(256), 0, st, p->sbeg.p, p->send.p, p->elem.p, p->mass.p, p->v.p, p->ke.p);
    if (!hip_ok(c, launched(), "kinetic launch")) return MLIP_ERR_DEVICE;
    hipLaunchKernelGGL(k_disp_next, dim3(nb), dim3(256), 0, st, p->disp.p, p->elem.p, p->mass.p, p->f.p, p->v.p,
                       p->x.p, p->xref.p, p->dnext.p);
    if (!hip_ok(c, launched(), "displacement launch") ||
        !hip_ok(c, hipMemcpyAsync(p->h_out, p->sout.p, 7 * nb * sizeof(double), hipMemcpyDeviceToHost, st),
                "pool results") ||
        !hip_ok(c, hipMemcpyAsync(p->h_out + 7 * nb, p->ke.p, 6 * nb * sizeof(double), hipMemcpyDeviceToHost, st),
                "pool kinetic") ||
        !hip_ok(c, hipMemcpyAsync(p->h_out + 13 * nb, p->dnext.p, nb * sizeof(double), hipMemcpyDeviceToHost, st),
                "pool displacement"))
        return MLIP_ERR_DEVICE;
    p->in_flight = true;
    return MLIP_OK;
}

int mlip_pool_sync(mlip_ctx* c, int* ncompleted) {
    if (!c || !c->pool) return MLIP_ERR_ARG;
    Pool* p = c->pool;
    if (ncompleted) *ncompleted = 0;
    if (!hip_ok(c, hipStreamSynchronize(c->stream), "pool sync")) return MLIP_ERR_DEVICE;
    if (!p->in_flight) return MLIP_OK;
    p->in_flight = false;
    if (c->prof) collect_profile(c);
    const int nb = static_cast<int>(p->batch.size());
    int ndone = 0;
    std::vector<int> rebuild;
    for (int b = 0; b < nb; ++b) {
        Request& r = p->req[p->batch[b]];
        for (int k = 0; k < 7; ++k) r.thermo.push_back(p->h_out[7 * b + k]);
        for (int k = 0; k < 6; ++k) r.thermo.push_back(p->h_out[7 * nb + 6 * b + k]);
        const bool last = r.nsteps == 0 || r.step == r.nsteps;
        if (!last) {
            ++r.step;
            if (p->h_out[13 * nb + b] > 0.5 * r.skin) rebuild.push_back(r.id);
            continue;
        }
        const size_t n3 = 3 * static_cast<size_t>(r.natoms);
        r.f.resize(n3);
        r.x.resize(n3);
        r.v.resize(n3);
        const size_t a3 = 3 * static_cast<size_t>(r.a0);
        if (!hip_ok(c, hipMemcpy(r.f.data(), p->f.p + a3, n3 * 8, hipMemcpyDeviceToHost), "pool forces") ||
            !hip_ok(c, hipMemcpy(r.x.data(), p->x.p + a3, n3 * 8, hipMemcpyDeviceToHost), "pool positions") ||
            !hip_ok(c, hipMemcpy(r.v.data(), p->v.p + a3, n3 * 8, hipMemcpyDeviceToHost), "pool velocities"))
            return MLIP_ERR_DEVICE;
        r.done = true;
        if (!release_request(c, p, r)) return MLIP_ERR_DEVICE;
        p->done_ids.push_back(r.id);
        ++ndone;
    }
    if (!rebuild.empty()) {
        const int rc = rebuild_lists(c, p, rebuild, &ndone);
        if (rc != MLIP_OK) return rc;
    }
    if (ncompleted) *ncompleted = ndone;
    return MLIP_OK;
}

int64_t mlip_pool_rebuilds(const mlip_ctx* c) { return c && c->pool ? c->pool->rebuilds : -1; }

int mlip_pool_done(mlip_ctx* c, int* ids, int cap) {
    if (!c || !c->pool || (!ids && cap > 0)) return -MLIP_ERR_ARG;
    Pool* p = c->pool;
    const int n = std::min<int>(cap, static_cast<int>(p->done_ids.size()));
    for (int k = 0; k < n; ++k) ids[k] = p->done_ids[k];
    p->done_ids.erase(p->done_ids.begin(), p->done_ids.begin() + n);
    return n;
}

int mlip_pool_result(mlip_ctx* c, int id, double* energy, double* virial, double* f, double* x, double* v,
                     double* thermo, int thermo_cap) {
    if (!c || !c->pool) return -MLIP_ERR_ARG;
    Pool* p = c->pool;
    auto it = p->req.find(id);
    if (it == p->req.end() || !it->second.done) {
        c->err = "request not completed";
        return -MLIP_ERR_STATE;
    }
    const Request& r = it->second;
    if (!r.error.empty()) {
        c->err = r.error;
        p->req.erase(it);
        return -MLIP_ERR_FULL;
    }
    const size_t nev = r.thermo.size() / kThermo;
    const double* lastv = r.thermo.data() + (nev - 1) * kThermo;
    if (energy) *energy = lastv[0];
    if (virial)
        for (int k = 0; k < 6; ++k) virial[k] = lastv[1 + k];
    if (f) std::copy(r.f.begin(), r.f.end(), f);
    if (x) std::copy(r.x.begin(), r.x.end(), x);
    if (v) std::copy(r.v.begin(), r.v.end(), v);
    if (thermo) std::copy(r.thermo.begin(), r.thermo.begin() + std::min<size_t>(r.thermo.size(), thermo_cap),
                          thermo);
    const int nret = static_cast<int>(nev);
    p->req.erase(it);
    return nret;
}

int mlip_pool_stats(const mlip_ctx* c, int64_t* out6) {
    if (!c || !c->pool || !out6) return MLIP_ERR_ARG;
    const Pool* p = c->pool;
    int64_t active = 0;
    for (const auto& kv : p->req) active += kv.second.done ? 0 : 1;
    out6[0] = p->atoms.used();
    out6[1] = p->pairs.used();
    out6[2] = p->cap_atoms;
    out6[3] = p->cap_pairs;
    out6[4] = static_cast<int64_t>(p->req.size());
    out6[5] = active;
    return MLIP_OK;
}

int64_t mlip_pool_bytes(const mlip_ctx* c, int64_t max_atoms, int64_t max_pairs, int max_requests) {
    if (!c || max_atoms < 1 || max_pairs < 1 || max_requests < 1) return -1;
    const Model& m = c->model;
    const int64_t slots = (max_atoms + static_cast<int64_t>(m.n_type) * kBlock + kBlock - 1) / kBlock * kBlock;
    const int64_t per_slot = 8 * (2 * static_cast<int64_t>(m.n_fn) * m.n_type * m.n_lmh * 2 + 3 * m.n_p + 3 +
                                  static_cast<int64_t>(m.n_fn) * (1 + kWavesPerWg)) +
                             8;
    const int64_t per_atom = 8 * (3 * 3 + 6) + 4 * 7;
    const int64_t per_pair = 4 + 8 * 3 + 8 * 3 + 2 * 4;
    return slots * per_slot + max_atoms * per_atom + max_pairs * per_pair + max_requests * (13 * 8 + 8 + 2 * 24);
}

}  // extern "C"
