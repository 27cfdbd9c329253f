// Binned periodic full neighbor list for owned atoms with image shifts.
//
// Builds what LAMMPS hands the Kokkos style with `neigh full newton on`
// (pair_mlip_gtinv_kokkos_impl.h:229-270), but for a single periodic cell
// without ghost atoms: neighbor j of atom i is an owned atom index plus the
// lattice shift s, del = x_i - x_j + s, every pair with 0 < |del| < rc.
// Self-images (j == i, s != 0) are kept, which small cells need.
// Reference for the binning idea: src/USER-MLIP/mlipkk_cell.cpp:218-307.
#include "neighbor.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

namespace mlip {

namespace {

void inv3(const double h[9], double g[9]) {
    // rows of h are a, b, c; g = inverse of the column matrix [a b c]
    const double a0 = h[0], a1 = h[1], a2 = h[2], b0 = h[3], b1 = h[4], b2 = h[5], c0 = h[6], c1 = h[7], c2 = h[8];
    // M columns = a, b, c  ->  M = [[a0 b0 c0],[a1 b1 c1],[a2 b2 c2]]
    const double det = a0 * (b1 * c2 - c1 * b2) - b0 * (a1 * c2 - c1 * a2) + c0 * (a1 * b2 - b1 * a2);
    const double id = 1.0 / det;
    g[0] = (b1 * c2 - c1 * b2) * id;
    g[1] = (c0 * b2 - b0 * c2) * id;
    g[2] = (b0 * c1 - c0 * b1) * id;
    g[3] = (c1 * a2 - a1 * c2) * id;
    g[4] = (a0 * c2 - c0 * a2) * id;
    g[5] = (c0 * a1 - a0 * c1) * id;
    g[6] = (a1 * b2 - b1 * a2) * id;
    g[7] = (b0 * a2 - a0 * b2) * id;
    g[8] = (a0 * b1 - b0 * a1) * id;
}

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

bool cell_bins(const double cell[9], double rc, CellBins& out, std::string& err) {
    if (!(rc > 0.0)) {
        err = "cutoff must be positive";
        return false;
    }
    inv3(cell, out.g);
    const double vol = std::fabs(cell[0] * (cell[4] * cell[8] - cell[5] * cell[7]) -
                                 cell[1] * (cell[3] * cell[8] - cell[5] * cell[6]) +
                                 cell[2] * (cell[3] * cell[7] - cell[4] * cell[6]));
    if (!(vol > 0.0) || !std::isfinite(vol)) {
        err = "degenerate cell";
        return false;
    }
    out.vol = vol;
    // face-to-face widths
    double w[3], cr[3];
    cross3(cell + 3, cell + 6, cr);
    w[0] = vol / norm3(cr);
    cross3(cell + 6, cell + 0, cr);
    w[1] = vol / norm3(cr);
    cross3(cell + 0, cell + 3, cr);
    w[2] = vol / norm3(cr);
    // bins along each lattice direction at least rc/2 wide (fewer distance
    // tests than rc-wide bins: a (5/2)^3 instead of 3^3 stencil volume);
    // the stencil reach in bins covers rc even with 1-2 bins
    int64_t nbins = 1;
    for (int k = 0; k < 3; ++k) {
        const double q = std::floor(2.0 * w[k] / rc);
        out.nb[k] = q >= 1.0 ? (q < 1 << 20 ? static_cast<int>(q) : 1 << 20) : 1;
        out.reach[k] = static_cast<int>(std::ceil(rc / (w[k] / out.nb[k])));
        nbins *= out.nb[k];
    }
    if (nbins > kMaxCellBins) {
        err = "cell too large for the bin grid (more than 2^28 bins)";
        return false;
    }
    return true;
}

bool build_full_list(const double cell[9], int n, const double* x, double rc, NeighborList& out, std::string& err,
                     int nthreads) {
    out.off.assign(static_cast<size_t>(n) + 1, 0);
    out.idx.clear();
    out.shift.clear();
    if (n == 0) return true;
    CellBins cb;
    if (!cell_bins(cell, rc, cb, err)) return false;
    for (int i = 0; i < 3 * n; ++i)
        if (!std::isfinite(x[i])) {
            err = "non-finite position";
            return false;
        }
    const double* g = cb.g;
    const double vol = cb.vol;
    // fractional coordinates wrapped into [0,1)
    std::vector<double> f(3 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const double* xi = x + 3 * i;
        for (int k = 0; k < 3; ++k) {
            double v = g[3 * k] * xi[0] + g[3 * k + 1] * xi[1] + g[3 * k + 2] * xi[2];
            v -= std::floor(v);
            if (v >= 1.0) v -= 1.0;
            f[3 * i + k] = v;
        }
    }
    const int* nb = cb.nb;
    const int nbins = nb[0] * nb[1] * nb[2];
    std::vector<int> bin_of(n), head(nbins + 1, 0), members(n);
    for (int i = 0; i < n; ++i) {
        int b[3];
        for (int k = 0; k < 3; ++k) b[k] = std::min(nb[k] - 1, static_cast<int>(f[3 * i + k] * nb[k]));
        bin_of[i] = (b[0] * nb[1] + b[1]) * nb[2] + b[2];
        ++head[bin_of[i] + 1];
    }
    for (int b = 0; b < nbins; ++b) head[b + 1] += head[b];
    {
        std::vector<int> fill(head.begin(), head.end() - 1);
        for (int i = 0; i < n; ++i) members[fill[bin_of[i]]++] = i;
    }
    const int* reach = cb.reach;
    const double rc2 = rc * rc;
    // positions reconstructed from wrapped fractional coordinates
    std::vector<double> xw(3 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            xw[3 * i + k] = f[3 * i] * cell[k] + f[3 * i + 1] * cell[3 + k] + f[3 * i + 2] * cell[6 + k];
    out.wrap_shift.assign(3 * static_cast<size_t>(n), 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) out.wrap_shift[3 * i + k] = x[3 * i + k] - xw[3 * i + k];
    // expected neighbors per atom, for reserving
    const double per_atom = 4.18879020478639 * rc * rc * rc * n / vol;
    // per bin, the stencil's neighbor bins and their lattice images, once
    const int sx = 2 * reach[0] + 1, sy = 2 * reach[1] + 1, sz = 2 * reach[2] + 1;
    const int ns = sx * sy * sz;
    std::vector<int> sbin(static_cast<size_t>(nbins) * ns);
    std::vector<double> ssh(3 * static_cast<size_t>(nbins) * ns);
    for (int bi = 0; bi < nbins; ++bi) {
        const int b0 = bi / (nb[1] * nb[2]), b1 = (bi / nb[2]) % nb[1], b2 = bi % nb[2];
        int t = 0;
        for (int d0 = -reach[0]; d0 <= reach[0]; ++d0)
            for (int d1 = -reach[1]; d1 <= reach[1]; ++d1)
                for (int d2 = -reach[2]; d2 <= reach[2]; ++d2, ++t) {
                    const int c0 = b0 + d0, c1 = b1 + d1, c2 = b2 + d2;
                    // bin index and the lattice image it lies in
                    const int j0 = ((c0 % nb[0]) + nb[0]) % nb[0], s0 = (c0 - j0) / nb[0];
                    const int j1 = ((c1 % nb[1]) + nb[1]) % nb[1], s1 = (c1 - j1) / nb[1];
                    const int j2 = ((c2 % nb[2]) + nb[2]) % nb[2], s2 = (c2 - j2) / nb[2];
                    const size_t e = static_cast<size_t>(bi) * ns + t;
                    sbin[e] = (j0 * nb[1] + j1) * nb[2] + j2;
                    for (int k = 0; k < 3; ++k) ssh[3 * e + k] = s0 * cell[k] + s1 * cell[3 + k] + s2 * cell[6 + k];
                }
    }
    int max_occ = 0;
    for (int b = 0; b < nbins; ++b) max_occ = std::max(max_occ, head[b + 1] - head[b]);
    const size_t max_cand = static_cast<size_t>(ns) * max_occ;
    auto scan = [&](int i0, int i1, std::vector<int32_t>& idx, std::vector<double>& shift, int64_t* cnt) {
        idx.reserve(static_cast<size_t>((i1 - i0) * per_atom * 1.15) + 16);
        shift.reserve(3 * idx.capacity());
        // branch-free candidate pass into a small buffer, then append hits
        std::vector<int32_t> cj(max_cand + 1);
        std::vector<double> cs(3 * (max_cand + 1));
        for (int i = i0; i < i1; ++i) {
            const size_t e0 = static_cast<size_t>(bin_of[i]) * ns;
            const double* xi = &xw[3 * i];
            const double* wi = &out.wrap_shift[3 * i];
            size_t k = 0;
            for (int t = 0; t < ns; ++t) {
                const int bj = sbin[e0 + t];
                const double* sv = &ssh[3 * (e0 + t)];
                const double ex = xi[0] - sv[0], ey = xi[1] - sv[1], ez = xi[2] - sv[2];
                const double ox = -sv[0] - wi[0], oy = -sv[1] - wi[1], oz = -sv[2] - wi[2];
                for (int q = head[bj]; q < head[bj + 1]; ++q) {
                    const int j = members[q];
                    // neighbor at x_j + S: del = x_i - x_j - S
                    const double dx = ex - xw[3 * j];
                    const double dy = ey - xw[3 * j + 1];
                    const double dz = ez - xw[3 * j + 2];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    // del = x_i - x_j + shift on the caller's (unwrapped) positions
                    const double* wj = &out.wrap_shift[3 * j];
                    cj[k] = j;
                    cs[3 * k] = ox + wj[0];
                    cs[3 * k + 1] = oy + wj[1];
                    cs[3 * k + 2] = oz + wj[2];
                    k += (r2 < rc2) & (r2 > 0.0);
                }
            }
            idx.insert(idx.end(), cj.begin(), cj.begin() + k);
            shift.insert(shift.end(), cs.begin(), cs.begin() + 3 * k);
            cnt[i] = static_cast<int64_t>(k);
        }
    };
    std::vector<int64_t> cnt(n);
    const int nt = std::max(1, std::min(nthreads, n / 256));
    if (nt == 1) {
        scan(0, n, out.idx, out.shift, cnt.data());
    } else {
        std::vector<std::vector<int32_t>> idx(nt);
        std::vector<std::vector<double>> sh(nt);
        std::vector<std::thread> th;
        auto range = [&](int t) { return std::make_pair(static_cast<int>(static_cast<int64_t>(n) * t / nt),
                                                        static_cast<int>(static_cast<int64_t>(n) * (t + 1) / nt)); };
        for (int t = 1; t < nt; ++t)
            th.emplace_back([&, t]() { scan(range(t).first, range(t).second, idx[t], sh[t], cnt.data()); });
        scan(range(0).first, range(0).second, idx[0], sh[0], cnt.data());
        for (auto& x_ : th) x_.join();
        size_t tot = 0;
        for (int t = 0; t < nt; ++t) tot += idx[t].size();
        out.idx.reserve(tot);
        out.shift.reserve(3 * tot);
        for (int t = 0; t < nt; ++t) {
            out.idx.insert(out.idx.end(), idx[t].begin(), idx[t].end());
            out.shift.insert(out.shift.end(), sh[t].begin(), sh[t].end());
        }
    }
    for (int i = 0; i < n; ++i) out.off[i + 1] = out.off[i] + cnt[i];
    return true;
}

}  // namespace mlip
