// Device-resident request pool: continuous batching of independent
// structures (single energy/force/stress evaluations and short NVE runs).
//
// Every admitted request owns an atom range and a pair range of the pool's
// device arrays (positions, velocities, element, per-centre pair ranges with
// image shifts, force-gather lists), filled once at admission from a
// neighbor list built with a skin on the device (the batch builder of
// nbuild.hip, on a stream of its own so that it runs beside a tick).  A tick evaluates all resident requests
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
#include <map>
#include <memory>
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

// A request's lists built in an admission batch (batch atom start A0, first
// pair P0) and its pool ranges (atoms a0.., pairs p0..).
struct AdmDesc {
    int64_t A0, P0, np, a0, p0;
    int32_t n, pad;
};

// Places admitted requests' device-built lists into their pool ranges:
// neighbor indices and signed gather entries (+-(p + 1)) move by the range
// offsets, image shifts are copied, per-atom pair and gather ranges come
// from the batch offsets.
__global__ void k_adm_place(const AdmDesc* d, const int32_t* __restrict__ bpoff, const int32_t* __restrict__ bgoff,
                            const int32_t* __restrict__ bpnbr, const double* __restrict__ bshift,
                            const int32_t* __restrict__ bgent, int32_t* pbeg, int32_t* pend, int32_t* gbeg,
                            int32_t* gend, int32_t* pnbr, double* pshift, int32_t* gent) {
    const AdmDesc q = d[blockIdx.y];
    const int64_t dp = q.p0 - q.P0;
    const int64_t m = q.n > 2 * q.np ? static_cast<int64_t>(q.n) : 2 * q.np;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride) {
        if (i < q.n) {
            const int64_t a = q.A0 + i, o = q.a0 + i;
            pbeg[o] = static_cast<int32_t>(bpoff[a] + dp);
            pend[o] = static_cast<int32_t>(bpoff[a + 1] + dp);
            gbeg[o] = static_cast<int32_t>(bgoff[a] + 2 * dp);
            gend[o] = static_cast<int32_t>(bgoff[a + 1] + 2 * dp);
        }
        if (i < q.np) {
            const int64_t e = q.P0 + i, o = q.p0 + i;
            pnbr[o] = static_cast<int32_t>(bpnbr[e] - q.A0 + q.a0);
            pshift[3 * o] = bshift[3 * e];
            pshift[3 * o + 1] = bshift[3 * e + 1];
            pshift[3 * o + 2] = bshift[3 * e + 2];
        }
        if (i < 2 * q.np) {
            const int32_t en = bgent[2 * q.P0 + i];
            gent[2 * q.p0 + i] = static_cast<int32_t>(en > 0 ? en + dp : en - dp);
        }
    }
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
    // admissions: lists built on the device on their own stream (while a
    // tick runs), then placed into the requests' ranges on the pool stream
    NbWork* nbw = nullptr;
    hipStream_t ast = nullptr;
    hipEvent_t built = nullptr;   // admission lists complete on the admission stream
    hipEvent_t placed = nullptr;  // last placement done (its inputs may be reused)
    bool placed_pending = false;
    DBuf<double> ax, apshift;
    DBuf<int32_t> apoff, apnbr, agoff, agent;
    DBuf<AdmDesc> adesc;
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
    if (p->ast) (void)hipStreamSynchronize(p->ast);
    nb_work_free(p->nbw);
    p->ax.release();
    p->apshift.release();
    p->apoff.release();
    p->apnbr.release();
    p->agoff.release();
    p->agent.release();
    p->adesc.release();
    if (p->placed) (void)hipEventDestroy(p->placed);
    if (p->built) (void)hipEventDestroy(p->built);
    if (p->ast) (void)hipStreamDestroy(p->ast);
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
    if (!ok || hipHostMalloc(reinterpret_cast<void**>(&p->h_out), 14 * nr * sizeof(double)) != hipSuccess ||
        hipStreamCreateWithFlags(&p->ast, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->placed, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->built, hipEventDisableTiming) != hipSuccess) {
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

// Admission with lists built on the device.  The requests 0..count-1 are
// checked in order (element mapping, finite positions, cell at cutoff +
// skin); their lists are built in one batch on the admission stream, which
// runs beside a tick in flight and is waited for alone to learn each
// request's pair count; ranges are then taken in FIFO order until one does
// not fit.  Positions, velocities, elements and reference positions go up
// through the pinned staging arena, and one kernel on the pool stream
// places the lists into the requests' ranges, ahead of the next tick.
// Returns the number admitted, or -MLIP_ERR_*.
int admit_device(mlip_ctx* c, Pool* p, int count, const int* natoms, const int64_t* at, const double* cells,
                 const double* x, const int* type, const double* v, const int* nsteps, double dt,
                 const double* skins, int* handles) {
    const double rcut = c->model.cutoff;
    std::vector<std::vector<int32_t>> elem(static_cast<size_t>(count));
    std::vector<double> rcs(static_cast<size_t>(count));
    int nc = 0;
    std::string first_err;
    for (; nc < count; ++nc) {
        std::string e;
        bool ok = true;
        elem[nc].resize(natoms[nc]);
        for (int i = 0; i < natoms[nc] && ok; ++i) {
            const int t = type[at[nc] + i];
            if (t < 1 || t > static_cast<int>(c->elem_of_type.size()) || c->elem_of_type[t - 1] < 0) {
                e = "atom type without an element mapping";
                ok = false;
            } else {
                elem[nc][i] = c->elem_of_type[t - 1];
            }
        }
        for (int64_t i = 3 * at[nc]; ok && i < 3 * at[nc + 1]; ++i)
            if (!std::isfinite(x[i])) {
                e = "non-finite position";
                ok = false;
            }
        rcs[nc] = rcut + skins[nc];
        CellBins cb;
        if (ok && !cell_bins(cells + 9 * static_cast<size_t>(nc), rcs[nc], cb, e)) ok = false;
        if (!ok) {
            first_err = "request structure: " + e;
            break;
        }
    }
    if (nc == 0) {
        c->err = first_err;
        return -MLIP_ERR_ARG;
    }
    nc = std::min<int>(nc, std::max(0, p->cap_req - static_cast<int>(p->req.size())));
    if (nc == 0) {
        c->err = "request pool: capacity exhausted";
        return 0;
    }
    // the admission buffers are free once the last placement has run
    if (p->placed_pending && !hip_ok(c, hipEventSynchronize(p->placed), "admission buffers")) return -MLIP_ERR_DEVICE;
    p->placed_pending = false;
    const size_t n3 = 3 * static_cast<size_t>(at[nc]);
    if (!p->ax.ensure(std::max<size_t>(n3, 1))) {
        c->err = "device allocation failed for admissions";
        return -MLIP_ERR_DEVICE;
    }
    if (!hip_ok(c, hipMemcpyAsync(p->ax.p, x, n3 * sizeof(double), hipMemcpyHostToDevice, p->ast), "admission positions"))
        return -MLIP_ERR_DEVICE;
    const NbOut out{&p->apoff, &p->apnbr, &p->apshift, &p->agoff, &p->agent};
    int64_t total = 0;
    std::vector<int64_t> spairs;
    if (!device_lists_into(c, p->nbw, p->ast, nc, natoms, cells, rcs.data(), p->ax.p, out, total, &spairs) ||
        !hip_ok(c, hipEventRecord(p->built, p->ast), "admission lists"))
        return -MLIP_ERR_DEVICE;
    // ranges in FIFO order
    std::vector<AdmDesc> dd;
    int n = 0;
    for (; n < nc; ++n) {
        const int64_t np = spairs[n + 1] - spairs[n];
        const int64_t a0 = p->atoms.take(natoms[n]);
        if (a0 < 0) break;
        const int64_t p0 = p->pairs.take(std::max<int64_t>(np, 1));
        if (p0 < 0) {
            p->atoms.give(a0, natoms[n]);
            break;
        }
        dd.push_back({at[n], spairs[n], np, a0, p0, natoms[n], 0});
    }
    if (n == 0) {
        c->err = "request pool: capacity exhausted";
        return 0;
    }
    // host data through the staging arena; the descriptors too
    if (!stage_begin(c, p)) return -MLIP_ERR_DEVICE;
    std::vector<double> zeros;
    for (int k = 0; k < n; ++k) {
        const size_t na = static_cast<size_t>(natoms[k]);
        const double* xk = x + 3 * at[k];
        const int64_t a0 = dd[k].a0;
        if (v == nullptr && zeros.size() < 3 * na) zeros.assign(3 * na, 0.0);
        if (!stage_upload(c, p, p->x.p + 3 * a0, xk, 3 * na, "positions") ||
            !stage_upload(c, p, p->xref.p + 3 * a0, xk, 3 * na, "reference positions") ||
            !stage_upload(c, p, p->v.p + 3 * a0, v ? v + 3 * at[k] : zeros.data(), 3 * na, "velocities") ||
            !stage_upload(c, p, p->elem.p + a0, elem[k].data(), na, "elements"))
            return -MLIP_ERR_DEVICE;
    }
    if (!p->adesc.ensure(dd.size())) {
        c->err = "device allocation failed for admissions";
        return -MLIP_ERR_DEVICE;
    }
    if (!stage_upload(c, p, p->adesc.p, dd.data(), dd.size(), "admission descriptors") || !stage_flush(c, p))
        return -MLIP_ERR_DEVICE;
    int64_t most = 1;
    for (const AdmDesc& q : dd) most = std::max<int64_t>(most, std::max<int64_t>(q.n, 2 * q.np));
    const unsigned gx = static_cast<unsigned>(std::min<int64_t>((most + 255) / 256, 64));
    if (!hip_ok(c, hipStreamWaitEvent(c->stream, p->built, 0), "admission lists")) return -MLIP_ERR_DEVICE;
    hipLaunchKernelGGL(k_adm_place, dim3(gx, static_cast<unsigned>(n)), dim3(256), 0, c->stream, p->adesc.p,
                       p->apoff.p, p->agoff.p, p->apnbr.p, p->apshift.p, p->agent.p, p->pbeg.p, p->pend.p,
                       p->gbeg.p, p->gend.p, p->pnbr.p, p->pshift.p, p->gent.p);
    if (!hip_ok(c, hipGetLastError(), "admission placement") ||
        !hip_ok(c, hipEventRecord(p->placed, c->stream), "admission placement"))
        return -MLIP_ERR_DEVICE;
    p->placed_pending = true;
    for (int k = 0; k < n; ++k) {
        Request r;
        r.natoms = natoms[k];
        r.nsteps = nsteps[k];
        r.dt = dt;
        r.skin = skins[k];
        std::copy(cells + 9 * k, cells + 9 * k + 9, r.cell);
        r.a0 = dd[k].a0;
        r.p0 = dd[k].p0;
        r.np = dd[k].np;
        r.elem = std::move(elem[k]);
        r.id = p->next_id++;
        handles[k] = r.id;
        p->req.emplace(r.id, std::move(r));
    }
    return n;
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
// ticks (the pool stream is idle), on the device: the requests' positions
// are gathered into the admission buffer, built as one batch, and placed
// into fresh pair ranges like an admission; the positions become the new
// reference positions.  A request whose new list does not fit the pair
// capacity fails with an error (reported by mlip_pool_result).
int rebuild_lists(mlip_ctx* c, Pool* p, const std::vector<int>& ids, int* ndone) {
    const int nr = static_cast<int>(ids.size());
    std::vector<Request*> rq(nr);
    std::vector<int32_t> natoms(nr);
    std::vector<int64_t> at(static_cast<size_t>(nr) + 1, 0);
    std::vector<double> cells(9 * static_cast<size_t>(nr)), rcs(nr);
    for (int k = 0; k < nr; ++k) {
        rq[k] = &p->req[ids[k]];
        natoms[k] = rq[k]->natoms;
        at[k + 1] = at[k] + natoms[k];
        std::copy(rq[k]->cell, rq[k]->cell + 9, cells.begin() + 9 * k);
        rcs[k] = c->model.cutoff + rq[k]->skin;
    }
    if (p->placed_pending && !hip_ok(c, hipEventSynchronize(p->placed), "rebuild buffers")) return MLIP_ERR_DEVICE;
    p->placed_pending = false;
    if (!p->ax.ensure(std::max<size_t>(3 * static_cast<size_t>(at[nr]), 1))) {
        c->err = "device allocation failed for list rebuilds";
        return MLIP_ERR_DEVICE;
    }
    for (int k = 0; k < nr; ++k)
        if (!hip_ok(c, hipMemcpyAsync(p->ax.p + 3 * at[k], p->x.p + 3 * rq[k]->a0, 24 * static_cast<size_t>(natoms[k]),
                                      hipMemcpyDeviceToDevice, c->stream),
                    "rebuild positions"))
            return MLIP_ERR_DEVICE;
    const NbOut out{&p->apoff, &p->apnbr, &p->apshift, &p->agoff, &p->agent};
    int64_t total = 0;
    std::vector<int64_t> spairs;
    if (!device_lists_into(c, p->nbw, c->stream, nr, natoms.data(), cells.data(), rcs.data(), p->ax.p, out, total,
                           &spairs))
        return MLIP_ERR_DEVICE;
    std::vector<AdmDesc> dd;
    for (int k = 0; k < nr; ++k) {
        Request& r = *rq[k];
        const int64_t np = spairs[k + 1] - spairs[k];
        p->pairs.give(r.p0, std::max<int64_t>(r.np, 1));
        const int64_t p0 = p->pairs.take(std::max<int64_t>(np, 1));
        if (p0 < 0) {
            // no list for the request: it ends here with an error
            r.error = "request pool: no pair capacity for an NVE neighbor-list rebuild";
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
        dd.push_back({at[k], spairs[k], np, r.a0, p0, r.natoms, 0});
        if (!hip_ok(c, hipMemcpyAsync(p->xref.p + 3 * r.a0, p->x.p + 3 * r.a0, 24 * static_cast<size_t>(r.natoms),
                                      hipMemcpyDeviceToDevice, c->stream),
                    "reference positions"))
            return MLIP_ERR_DEVICE;
    }
    if (dd.empty()) return MLIP_OK;
    if (!p->adesc.ensure(dd.size())) {
        c->err = "device allocation failed for list rebuilds";
        return MLIP_ERR_DEVICE;
    }
    if (!stage_begin(c, p) || !stage_upload(c, p, p->adesc.p, dd.data(), dd.size(), "rebuild descriptors") ||
        !stage_flush(c, p))
        return MLIP_ERR_DEVICE;
    int64_t most = 1;
    for (const AdmDesc& q : dd) most = std::max<int64_t>(most, std::max<int64_t>(q.n, 2 * q.np));
    const unsigned gx = static_cast<unsigned>(std::min<int64_t>((most + 255) / 256, 64));
    hipLaunchKernelGGL(k_adm_place, dim3(gx, static_cast<unsigned>(dd.size())), dim3(256), 0, c->stream, p->adesc.p,
                       p->apoff.p, p->agoff.p, p->apnbr.p, p->apshift.p, p->agent.p, p->pbeg.p, p->pend.p,
                       p->gbeg.p, p->gend.p, p->pnbr.p, p->pshift.p, p->gent.p);
    if (!hip_ok(c, hipGetLastError(), "rebuild placement") ||
        !hip_ok(c, hipEventRecord(p->placed, c->stream), "rebuild placement"))
        return MLIP_ERR_DEVICE;
    p->placed_pending = true;
    return MLIP_OK;
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
    const int64_t at[2] = {0, natoms};
    const int k = admit_device(c, c->pool, 1, &natoms, at, cell, x, type, v, &nsteps, dt, &skin, handle);
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
    (void)nthreads;  // the lists are built on the device
    return admit_device(c, c->pool, count, natoms, at.data(), cells, x, type, v, nsteps, dt, skin, handles);
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
