// Device-side periodic neighbor lists for one structure or a batch of
// independent structures, and the force gather lists built from them.
//
// The lists are the same, entry for entry and bit for bit, as the host
// builder's (neighbor.cpp): the same bin grid (cell_bins), atoms wrapped
// into the cell with the same double operations (this file is compiled
// without floating-point contraction, like the host build), bin members in
// increasing atom order (stable radix sort), the stencil walked in the same
// order, and pairs kept with 0 < r^2 < rc^2 on the same reconstructed
// positions.  So a structure evaluated from device-built lists matches one
// evaluated from host-built lists exactly.
//
// Reference for what a full list holds: LAMMPS `neigh full newton on` as
// requested by pair_mlip_gtinv_kokkos_impl.h:229-270, and the binned
// periodic search of src/USER-MLIP/mlipkk_cell.cpp:218-307 (here with
// self-images and the strict r < rc test of the pair style).
//
// Passes (thread per atom unless noted):
//   k_nb_bin     fractional coordinates, wrapped position, global bin key
//   radix sort   (bin key, atom) -> bin members in atom order
//   k_nb_heads   thread per bin: first member (binary search)
//   k_nb_pairs   count pass, exclusive scan, fill pass
//   radix sort   (neighbor, pair) -> incoming pairs of every position in
//                pair order
//   k_nb_gather  thread per position: merge its outgoing (-) and incoming
//                (+) pairs in pair order
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstring>

#include "engine_internal.h"
#include "neighbor.h"

#pragma clang fp contract(off)

namespace mlip {

struct NbCell {
    double cell[9], g[9];
    double rc2;  // squared cutoff of the structure
    int32_t nb[3], reach[3];
    int32_t bin0, natoms;
};

struct NbWork {
    DBuf<NbCell> cells;
    DBuf<int32_t> sid, val, val2, head, pv, pv2, ihead;
    DBuf<uint32_t> key, key2, pk, pk2;
    DBuf<double> xw, wrap;
    DBuf<int64_t> cnt, off, first, soff;
    DBuf<char> tmp;
    void release() {
        first.release();
        soff.release();
        cells.release();
        sid.release();
        val.release();
        val2.release();
        head.release();
        pv.release();
        pv2.release();
        ihead.release();
        key.release();
        key2.release();
        pk.release();
        pk2.release();
        xw.release();
        wrap.release();
        cnt.release();
        off.release();
        tmp.release();
    }
};

namespace {

constexpr int kNbThreads = 256;

inline unsigned grid_of(int64_t n) { return static_cast<unsigned>((n + kNbThreads - 1) / kNbThreads); }

__global__ __launch_bounds__(kNbThreads) void k_nb_bin(const NbCell* __restrict__ cells,
                                                       const int32_t* __restrict__ sid, int n,
                                                       const double* __restrict__ x, double* __restrict__ xw,
                                                       double* __restrict__ wrap, uint32_t* __restrict__ key,
                                                       int32_t* __restrict__ val) {
    const int a = blockIdx.x * kNbThreads + threadIdx.x;
    if (a >= n) return;
    const NbCell& C = cells[sid[a]];
    const double x0 = x[3 * a], x1 = x[3 * a + 1], x2 = x[3 * a + 2];
    double f[3];
    int b[3];
    for (int k = 0; k < 3; ++k) {
        double v = C.g[3 * k] * x0 + C.g[3 * k + 1] * x1 + C.g[3 * k + 2] * x2;
        v -= floor(v);
        if (v >= 1.0) v -= 1.0;
        f[k] = v;
        b[k] = max(0, min(C.nb[k] - 1, static_cast<int>(v * C.nb[k])));
    }
    for (int k = 0; k < 3; ++k) {
        const double w = f[0] * C.cell[k] + f[1] * C.cell[3 + k] + f[2] * C.cell[6 + k];
        xw[3 * a + k] = w;
        wrap[3 * a + k] = x[3 * a + k] - w;
    }
    key[a] = static_cast<uint32_t>(C.bin0 + (b[0] * C.nb[1] + b[1]) * C.nb[2] + b[2]);
    val[a] = a;
}

// head[b] = first index of sorted[] holding a key >= b, for b in [0, nkeys].
__global__ __launch_bounds__(kNbThreads) void k_nb_heads(const uint32_t* __restrict__ sorted, int n, int nkeys,
                                                         int32_t* __restrict__ head) {
    const int b = blockIdx.x * kNbThreads + threadIdx.x;
    if (b > nkeys) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sorted[mid] < static_cast<uint32_t>(b))
            lo = mid + 1;
        else
            hi = mid;
    }
    head[b] = lo;
}

// Count (FILL = false) or write (FILL = true) atom a's neighbors: stencil
// bins in (d0, d1, d2) order, members in atom order, 0 < r^2 < rc2.
template <bool FILL>
__global__ __launch_bounds__(kNbThreads) void k_nb_pairs(const NbCell* __restrict__ cells,
                                                         const int32_t* __restrict__ sid, int n,
                                                         const uint32_t* __restrict__ key,
                                                         const int32_t* __restrict__ head,
                                                         const int32_t* __restrict__ member,
                                                         const double* __restrict__ xw,
                                                         const double* __restrict__ wrap,
                                                         int64_t* __restrict__ cnt, const int64_t* __restrict__ off,
                                                         int32_t* __restrict__ pnbr, double* __restrict__ pshift) {
    const int a = blockIdx.x * kNbThreads + threadIdx.x;
    if (a >= n) return;
    const NbCell& C = cells[sid[a]];
    const double rc2 = C.rc2;
    const int nb0 = C.nb[0], nb1 = C.nb[1], nb2 = C.nb[2];
    const int bl = static_cast<int>(key[a]) - C.bin0;
    const int b0 = bl / (nb1 * nb2), b1 = (bl / nb2) % nb1, b2 = bl % nb2;
    const double xi0 = xw[3 * a], xi1 = xw[3 * a + 1], xi2 = xw[3 * a + 2];
    const double wi0 = wrap[3 * a], wi1 = wrap[3 * a + 1], wi2 = wrap[3 * a + 2];
    int64_t o = 0, lim = 0;
    if (FILL) {
        o = off[a];
        lim = off[a + 1] - o;
    }
    int64_t k = 0;
    for (int d0 = -C.reach[0]; d0 <= C.reach[0]; ++d0) {
        const int c0 = b0 + d0;
        const int j0 = ((c0 % nb0) + nb0) % nb0, s0 = (c0 - j0) / nb0;
        for (int d1 = -C.reach[1]; d1 <= C.reach[1]; ++d1) {
            const int c1 = b1 + d1;
            const int j1 = ((c1 % nb1) + nb1) % nb1, s1 = (c1 - j1) / nb1;
            for (int d2 = -C.reach[2]; d2 <= C.reach[2]; ++d2) {
                const int c2 = b2 + d2;
                const int j2 = ((c2 % nb2) + nb2) % nb2, s2 = (c2 - j2) / nb2;
                const int bj = C.bin0 + (j0 * nb1 + j1) * nb2 + j2;
                // lattice image of the stencil bin: the neighbor sits at x_j + S
                const double sv0 = static_cast<double>(s0) * C.cell[0] + static_cast<double>(s1) * C.cell[3] +
                                   static_cast<double>(s2) * C.cell[6];
                const double sv1 = static_cast<double>(s0) * C.cell[1] + static_cast<double>(s1) * C.cell[4] +
                                   static_cast<double>(s2) * C.cell[7];
                const double sv2 = static_cast<double>(s0) * C.cell[2] + static_cast<double>(s1) * C.cell[5] +
                                   static_cast<double>(s2) * C.cell[8];
                const double ex = xi0 - sv0, ey = xi1 - sv1, ez = xi2 - sv2;
                const double ox = -sv0 - wi0, oy = -sv1 - wi1, oz = -sv2 - wi2;
                const int q1 = head[bj + 1];
                for (int q = head[bj]; q < q1; ++q) {
                    const int j = member[q];
                    const double dx = ex - xw[3 * j];
                    const double dy = ey - xw[3 * j + 1];
                    const double dz = ez - xw[3 * j + 2];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 < rc2 && r2 > 0.0) {
                        if (FILL && k < lim) {
                            // del = x_i - x_j + shift on the caller's positions
                            const int64_t e = o + k;
                            pnbr[e] = j;
                            pshift[3 * e] = ox + wrap[3 * j];
                            pshift[3 * e + 1] = oy + wrap[3 * j + 1];
                            pshift[3 * e + 2] = oz + wrap[3 * j + 2];
                        }
                        ++k;
                    }
                }
            }
        }
    }
    if (!FILL) cnt[a] = k;
}

__global__ __launch_bounds__(kNbThreads) void k_nb_off32(const int64_t* __restrict__ off, int n,
                                                         int32_t* __restrict__ poff) {
    const int a = blockIdx.x * kNbThreads + threadIdx.x;
    if (a <= n) poff[a] = static_cast<int32_t>(off[a]);
}

__global__ __launch_bounds__(kNbThreads) void k_nb_pair_keys(const int32_t* __restrict__ pnbr, int np,
                                                             uint32_t* __restrict__ pk, int32_t* __restrict__ pv) {
    const int p = blockIdx.x * kNbThreads + threadIdx.x;
    if (p >= np) return;
    pk[p] = static_cast<uint32_t>(pnbr[p]);
    pv[p] = p;
}

// first pair of each structure: off at the structure's first atom
__global__ __launch_bounds__(kNbThreads) void k_nb_struct_off(const int64_t* __restrict__ off,
                                                              const int64_t* __restrict__ first_atom, int nstruct,
                                                              int64_t* __restrict__ out) {
    const int s = blockIdx.x * kNbThreads + threadIdx.x;
    if (s <= nstruct) out[s] = off[first_atom[s]];
}

// Gather list of position k: its own pairs (centre k, entry -(p+1)) and the
// pairs that have k as neighbor (entry p+1), in increasing p; for a
// self-image pair (both), the centre entry first.
__global__ __launch_bounds__(kNbThreads) void k_nb_gather(const int32_t* __restrict__ poff,
                                                          const int32_t* __restrict__ ihead,
                                                          const int32_t* __restrict__ inc, int n,
                                                          int32_t* __restrict__ goff, int32_t* __restrict__ gent) {
    const int k = blockIdx.x * kNbThreads + threadIdx.x;
    if (k > n) return;
    const int w0 = poff[k] + ihead[k];
    goff[k] = w0;
    if (k == n) return;
    int po = poff[k];
    const int pe = poff[k + 1];
    int q = ihead[k];
    const int qe = ihead[k + 1];
    int w = w0;
    while (po < pe || q < qe) {
        if (q >= qe || (po < pe && po <= inc[q])) {
            gent[w++] = -(po + 1);
            ++po;
        } else {
            gent[w++] = inc[q] + 1;
            ++q;
        }
    }
}

int key_bits(int64_t nkeys) {
    int b = 1;
    while (b < 32 && (int64_t(1) << b) < nkeys) ++b;
    return b;
}

bool cub_ok(mlip_ctx* c, hipError_t e, const char* what) { return hip_ok(c, e, what); }

}  // namespace

bool device_lists_into(mlip_ctx* c, NbWork*& wp, hipStream_t st, int nstruct, const int32_t* natoms,
                       const double* cells, const double* rcs, const double* x, const NbOut& out, int64_t& npairs,
                       std::vector<int64_t>* spairs) {
    if (!wp) wp = new NbWork();
    NbWork& w = *wp;
    auto up = [&](auto& b, const auto* h, size_t n, const char* what) {
        if (!b.ensure(n)) {
            c->err = std::string("device allocation failed for ") + what;
            return false;
        }
        return n == 0 || hip_ok(c, hipMemcpyAsync(b.p, h, n * sizeof(*h), hipMemcpyHostToDevice, st), what);
    };
    npairs = 0;
    std::vector<NbCell> hc(static_cast<size_t>(nstruct));
    std::vector<int32_t> sid;
    std::vector<int64_t> first(static_cast<size_t>(nstruct) + 1, 0);
    int64_t bins = 0, n64 = 0;
    for (int s = 0; s < nstruct; ++s) {
        first[s] = n64;
        n64 += natoms[s];
    }
    first[nstruct] = n64;
    if (n64 > INT_MAX - 1) {
        c->err = "too many atoms";
        return false;
    }
    sid.reserve(static_cast<size_t>(n64));
    for (int s = 0; s < nstruct; ++s) {
        NbCell& h = hc[s];
        std::memset(&h, 0, sizeof h);
        h.bin0 = static_cast<int32_t>(bins);
        h.natoms = natoms[s];
        h.rc2 = rcs[s] * rcs[s];
        if (natoms[s] == 0) continue;
        CellBins cb;
        std::string e;
        if (!cell_bins(cells + 9 * static_cast<size_t>(s), rcs[s], cb, e)) {
            c->err = "structure " + std::to_string(s) + ": " + e;
            return false;
        }
        for (int k = 0; k < 9; ++k) {
            h.cell[k] = cells[9 * static_cast<size_t>(s) + k];
            h.g[k] = cb.g[k];
        }
        for (int k = 0; k < 3; ++k) {
            h.nb[k] = cb.nb[k];
            h.reach[k] = cb.reach[k];
        }
        bins += static_cast<int64_t>(cb.nb[0]) * cb.nb[1] * cb.nb[2];
        if (bins > kMaxCellBins) {
            c->err = "batch too large for the bin grid (more than 2^28 bins)";
            return false;
        }
        sid.insert(sid.end(), natoms[s], s);
    }
    const int n = static_cast<int>(n64);
    if (!out.poff->ensure(static_cast<size_t>(n) + 1) || !out.goff->ensure(static_cast<size_t>(n) + 1)) {
        c->err = "device allocation failed for neighbor offsets";
        return false;
    }
    if (n == 0) {
        if (spairs) spairs->assign(static_cast<size_t>(nstruct) + 1, 0);
        return hip_ok(c, hipMemsetAsync(out.poff->p, 0, sizeof(int32_t), st), "neighbor offsets") &&
               hip_ok(c, hipMemsetAsync(out.goff->p, 0, sizeof(int32_t), st), "gather offsets") &&
               hip_ok(c, hipStreamSynchronize(st), "neighbor offsets");
    }
    const int nbins = static_cast<int>(bins);
    if (!up(w.cells, hc.data(), hc.size(), "bin grids") || !up(w.sid, sid.data(), sid.size(), "atoms") ||
        !up(w.first, first.data(), first.size(), "structure starts"))
        return false;
    if (!w.xw.ensure(3 * static_cast<size_t>(n)) || !w.wrap.ensure(3 * static_cast<size_t>(n)) ||
        !w.key.ensure(n) || !w.key2.ensure(n) || !w.val.ensure(n) || !w.val2.ensure(n) ||
        !w.head.ensure(static_cast<size_t>(nbins) + 1) || !w.cnt.ensure(static_cast<size_t>(n) + 1) ||
        !w.off.ensure(static_cast<size_t>(n) + 1) || !w.soff.ensure(static_cast<size_t>(nstruct) + 1)) {
        c->err = "device allocation failed for the neighbor build";
        return false;
    }
    // scratch for the sorts and the scan
    size_t t_bins = 0, t_scan = 0;
    if (!cub_ok(c, hipcub::DeviceRadixSort::SortPairs(nullptr, t_bins, w.key.p, w.key2.p, w.val.p, w.val2.p, n, 0,
                                                      key_bits(nbins), st),
                "bin sort size") ||
        !cub_ok(c, hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan, w.cnt.p, w.off.p, n + 1, st), "scan size"))
        return false;
    if (!w.tmp.ensure(std::max(t_bins, t_scan))) {
        c->err = "device allocation failed for the neighbor build";
        return false;
    }
    k_nb_bin<<<grid_of(n), kNbThreads, 0, st>>>(w.cells.p, w.sid.p, n, x, w.xw.p, w.wrap.p, w.key.p, w.val.p);
    size_t tb = w.tmp.cap;
    if (!cub_ok(c, hipcub::DeviceRadixSort::SortPairs(w.tmp.p, tb, w.key.p, w.key2.p, w.val.p, w.val2.p, n, 0,
                                                      key_bits(nbins), st),
                "bin sort"))
        return false;
    k_nb_heads<<<grid_of(static_cast<int64_t>(nbins) + 1), kNbThreads, 0, st>>>(w.key2.p, n, nbins, w.head.p);
    k_nb_pairs<false><<<grid_of(n), kNbThreads, 0, st>>>(w.cells.p, w.sid.p, n, w.key.p, w.head.p, w.val2.p, w.xw.p,
                                                         w.wrap.p, w.cnt.p, nullptr, nullptr, nullptr);
    if (!hip_ok(c, hipMemsetAsync(w.cnt.p + n, 0, sizeof(int64_t), st), "neighbor counts")) return false;
    tb = w.tmp.cap;
    if (!cub_ok(c, hipcub::DeviceScan::ExclusiveSum(w.tmp.p, tb, w.cnt.p, w.off.p, n + 1, st), "neighbor scan"))
        return false;
    std::vector<int64_t> soff(static_cast<size_t>(nstruct) + 1, 0);
    k_nb_struct_off<<<grid_of(static_cast<int64_t>(nstruct) + 1), kNbThreads, 0, st>>>(w.off.p, w.first.p, nstruct,
                                                                                       w.soff.p);
    if (!hip_ok(c, hipMemcpyAsync(soff.data(), w.soff.p, soff.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st),
                "pair count") ||
        !hip_ok(c, hipStreamSynchronize(st), "pair count"))
        return false;
    const int64_t total = soff[nstruct];
    if (spairs) *spairs = soff;
    // gather offsets reach 2 * npairs and are int32
    if (total > (INT_MAX - 1) / 2) {
        c->err = "more than 2^30 pairs";
        return false;
    }
    const int np = static_cast<int>(total);
    npairs = total;
    if (!out.pnbr->ensure(np) || !out.pshift->ensure(3 * static_cast<size_t>(np)) ||
        !out.gent->ensure(2 * static_cast<size_t>(np)) || !w.ihead.ensure(static_cast<size_t>(n) + 1)) {
        c->err = "device allocation failed for the neighbor list";
        return false;
    }
    k_nb_pairs<true><<<grid_of(n), kNbThreads, 0, st>>>(w.cells.p, w.sid.p, n, w.key.p, w.head.p, w.val2.p, w.xw.p,
                                                        w.wrap.p, nullptr, w.off.p, out.pnbr->p, out.pshift->p);
    k_nb_off32<<<grid_of(static_cast<int64_t>(n) + 1), kNbThreads, 0, st>>>(w.off.p, n, out.poff->p);
    if (np == 0) {
        return hip_ok(c, hipMemsetAsync(out.goff->p, 0, (static_cast<size_t>(n) + 1) * sizeof(int32_t), st),
                      "gather offsets");
    }
    // incoming pairs of every position, in pair order
    if (!w.pk.ensure(np) || !w.pk2.ensure(np) || !w.pv.ensure(np) || !w.pv2.ensure(np)) {
        c->err = "device allocation failed for the gather build";
        return false;
    }
    size_t t_pairs = 0;
    if (!cub_ok(c, hipcub::DeviceRadixSort::SortPairs(nullptr, t_pairs, w.pk.p, w.pk2.p, w.pv.p, w.pv2.p, np, 0,
                                                      key_bits(n), st),
                "pair sort size"))
        return false;
    if (!w.tmp.ensure(t_pairs)) {
        c->err = "device allocation failed for the gather build";
        return false;
    }
    k_nb_pair_keys<<<grid_of(np), kNbThreads, 0, st>>>(out.pnbr->p, np, w.pk.p, w.pv.p);
    tb = w.tmp.cap;
    if (!cub_ok(c, hipcub::DeviceRadixSort::SortPairs(w.tmp.p, tb, w.pk.p, w.pk2.p, w.pv.p, w.pv2.p, np, 0,
                                                      key_bits(n), st),
                "pair sort"))
        return false;
    k_nb_heads<<<grid_of(static_cast<int64_t>(n) + 1), kNbThreads, 0, st>>>(w.pk2.p, np, n, w.ihead.p);
    k_nb_gather<<<grid_of(static_cast<int64_t>(n) + 1), kNbThreads, 0, st>>>(out.poff->p, w.ihead.p, w.pv2.p, n,
                                                                             out.goff->p, out.gent->p);
    return hip_ok(c, hipGetLastError(), "neighbor build");
}

bool device_lists(mlip_ctx* c, int nstruct, const int32_t* natoms, const double* cells, double rc,
                  int64_t& npairs) {
    const std::vector<double> rcs(static_cast<size_t>(nstruct), rc);
    const NbOut out{&c->poff, &c->pnbr, &c->pshift, &c->goff, &c->gent};
    return device_lists_into(c, c->nbw, c->stream, nstruct, natoms, cells, rcs.data(), c->x.p, out, npairs, nullptr);
}

void nb_work_free(NbWork*& w) {
    if (!w) return;
    w->release();
    delete w;
    w = nullptr;
}

void nb_free(mlip_ctx* c) { nb_work_free(c->nbw); }

}  // namespace mlip
