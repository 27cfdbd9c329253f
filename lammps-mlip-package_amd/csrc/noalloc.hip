// mlip_compute_flat_noalloc: the LAMMPS-shaped evaluation with every device
// buffer taken from one file-scope device array instead of hipMalloc, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid device allocation).
//
// The array holds kScratchBytes per GPU for the life of the process; its
// device address is resolved once per device (scratch_setup, on the entry
// point's first call on that device, with the context's device current) and
// only that address is used by host code.  Every call on a GPU reuses the same array and leaves
// the last call's contents in it: calls that overlap on one GPU (two
// contexts, two host threads) race.  What the kernels expect to start at
// zero (a_nlm, the per-centre virial) is cleared by a fill kernel each call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "engine_internal.h"

namespace mlip {

namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(128) << 20;  // 128 MiB per GPU
constexpr size_t kScratchAlign = 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // per device

// Sub-allocation from the scratch array; a pass with base == nullptr only
// measures.
struct Carver {
    unsigned char* base;
    size_t used = 0;
    template <typename T>
    T* take(size_t n) {
        used = (used + kScratchAlign - 1) / kScratchAlign * kScratchAlign;
        T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
        used += std::max<size_t>(n, 1) * sizeof(T);
        return p;
    }
};

struct Plan {
    SysBufs b{};
    double *epv = nullptr, *evat = nullptr;
};

void carve(Carver& cv, const Model& m, const LammpsLists& L, int inum, const WorkSizes& z, bool split, Plan& P) {
    const size_t np = static_cast<size_t>(L.npairs), na = static_cast<size_t>(L.nall);
    SysBufs& b = P.b;
    b.x = cv.take<double>(3 * na);
    b.pshift = nullptr;
    b.elem = cv.take<int32_t>(na);
    b.poff = cv.take<int32_t>(static_cast<size_t>(inum) + 1);
    b.pnbr = cv.take<int32_t>(np);
    b.pcen = cv.take<int32_t>(np);
    b.goff = cv.take<int32_t>(na + 1);
    b.gent = cv.take<int32_t>(2 * np);
    b.atom_of_slot = cv.take<int32_t>(z.nslot);
    b.slot_of_atom = cv.take<int32_t>(static_cast<size_t>(inum));
    b.block_type = cv.take<int32_t>(z.nblocks);
    b.A = cv.take<double>(z.asz);
    b.G = cv.take<double>(z.asz);
    b.D = cv.take<double>(z.psz);
    b.Hq = cv.take<double>(z.psz);
    b.H = cv.take<double>(z.psz);
    b.Ep = cv.take<double>(z.nslot);
    b.Eq12 = cv.take<double>(z.nslot * m.n_fn);
    b.Eq34 = cv.take<double>(z.nslot * m.n_fn * kWavesPerWg);
    b.gpair = cv.take<double>(3 * np);
    b.vown = cv.take<double>(6 * static_cast<size_t>(inum));
    b.f = cv.take<double>(3 * na);
    b.e_slot = cv.take<double>(z.nslot);
    b.red = cv.take<double>(7 * 128);
    b.out = cv.take<double>(8);
    if (split) {
        P.epv = cv.take<double>(7 * np);
        P.evat = cv.take<double>(7 * na);
    }
}

// Device address of this device's copy of the array (resolved on the first
// call per device; `device` must be current).
unsigned char* scratch_setup(int device, std::string& err) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0) {
        err = "no device";
        return nullptr;
    }
    if (static_cast<int>(g_scratch_ptr.size()) <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return g_scratch_ptr[device];
    // the symbol resolves to the copy of the device that is current
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) {
        err = "no current device";
        return nullptr;
    }
    if (cur != device && hipSetDevice(device) != hipSuccess) {
        err = "cannot make the context's device current";
        return nullptr;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
        err = std::string("scratch array address: ") + hipGetErrorString(e);
        return nullptr;
    }
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    return g_scratch_ptr[device];
}

}  // namespace

}  // namespace mlip

using namespace mlip;

extern "C" int mlip_compute_flat_noalloc(mlip_ctx* c, int nlocal, int nghost, const double* x, const int* type,
                                         const int64_t* tag, int inum, const int* ilist, const int* numneigh,
                                         const int* neigh_off, const int* neigh_idx, int neigh_full, int eflag,
                                         int vflag, double* f, double* eng, double* virial, double* eatom,
                                         double* vatom) {
    (void)eflag;
    (void)vflag;
    if (!c || !neigh_off || !neigh_idx || !f) return MLIP_ERR_ARG;
    std::string serr;
    unsigned char* base = scratch_setup(c->device, serr);
    if (!base) {
        c->err = "no scratch array on this device: " + serr;
        return MLIP_ERR_DEVICE;
    }
    LammpsLists L;
    int rc = lammps_lists(c, nlocal, nghost, x, type, tag, inum, ilist, numneigh, nullptr, neigh_off, neigh_idx,
                          neigh_full, L);
    if (rc != MLIP_OK) return rc;
    const Model& m = c->model;
    std::vector<int32_t> aos, sos, bt;
    assign_slots(c, inum, L.elem_own, aos, sos, bt);
    const WorkSizes z = work_sizes(m, L.nall, inum, L.npairs, static_cast<int>(bt.size()));
    const bool split = (eatom || vatom) && c->atom_mode == MLIP_ATOM_PAIR_SPLIT;
    Plan P;
    Carver measure{nullptr};
    carve(measure, m, L, inum, z, split, P);
    if (measure.used > kScratchBytes) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "structure needs %.1f MiB of device scratch; the no-allocation entry point holds %zu MiB per "
                      "GPU (use mlip_compute_flat)",
                      measure.used / 1048576.0, kScratchBytes >> 20);
        c->err = buf;
        return MLIP_ERR_ARG;
    }
    Carver cv{base};
    carve(cv, m, L, inum, z, split, P);
    hipStream_t st = c->stream;
    const size_t np = static_cast<size_t>(L.npairs), na = static_cast<size_t>(L.nall);
    auto up = [&](void* dst, const void* src, size_t bytes, const char* what) {
        return bytes == 0 || hip_ok(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), what);
    };
    if (!up(P.b.x, L.x.data(), 3 * na * sizeof(double), "positions") ||
        !up(P.b.elem, L.elem.data(), na * sizeof(int32_t), "elements") ||
        !up(P.b.poff, L.poff.data(), L.poff.size() * sizeof(int32_t), "pair offsets") ||
        !up(P.b.pnbr, L.pnbr.data(), np * sizeof(int32_t), "pair neighbors") ||
        !up(P.b.pcen, L.pcen.data(), np * sizeof(int32_t), "pair centres") ||
        !up(P.b.goff, L.gof.data(), L.gof.size() * sizeof(int32_t), "gather offsets") ||
        !up(P.b.gent, L.gen.data(), L.gen.size() * sizeof(int32_t), "gather entries"))
        return MLIP_ERR_DEVICE;
    c->resident = false;
    c->nstruct = 0;
    if (!bind_system(c, L.nall, inum, L.npairs, L.elem_own, aos, sos, bt, P.b)) return MLIP_ERR_DEVICE;
    if (!run_pipeline(c)) return MLIP_ERR_DEVICE;
    return lammps_results(c, L.nall, inum, ilist, f, eng, virial, eatom, vatom, P.epv, P.evat);
}
