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

bool build_full_list(const double cell[9], int n, const double* x, double rc, NeighborList& out, std::string& err) {
    out.off.assign(static_cast<size_t>(n) + 1, 0);
    out.idx.clear();
    out.shift.clear();
    if (n == 0) return true;
    if (!(rc > 0.0)) {
        err = "cutoff must be positive";
        return false;
    }
    double g[9];
    inv3(cell, g);
    const double vol = std::fabs(cell[0] * (cell[4] * cell[8] - cell[5] * cell[7]) -
                                 cell[1] * (cell[3] * cell[8] - cell[5] * cell[6]) +
                                 cell[2] * (cell[3] * cell[7] - cell[4] * cell[6]));
    if (!(vol > 0.0)) {
        err = "degenerate cell";
        return false;
    }
    // face-to-face widths
    double w[3], cr[3];
    cross3(cell + 3, cell + 6, cr);
    w[0] = vol / norm3(cr);
    cross3(cell + 6, cell + 0, cr);
    w[1] = vol / norm3(cr);
    cross3(cell + 0, cell + 3, cr);
    w[2] = vol / norm3(cr);
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
    // bins along each lattice direction at least rc wide
    int nb[3];
    for (int k = 0; k < 3; ++k) nb[k] = std::max(1, static_cast<int>(std::floor(w[k] / rc)));
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
    // stencil reach in bins per direction (covers rc even with 1-2 bins)
    int reach[3];
    for (int k = 0; k < 3; ++k) reach[k] = static_cast<int>(std::ceil(rc / (w[k] / nb[k])));
    const double rc2 = rc * rc;
    // positions reconstructed from wrapped fractional coordinates
    std::vector<double> xw(3 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            xw[3 * i + k] = f[3 * i] * cell[k] + f[3 * i + 1] * cell[3 + k] + f[3 * i + 2] * cell[6 + k];
    out.wrap_shift.assign(3 * static_cast<size_t>(n), 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) out.wrap_shift[3 * i + k] = x[3 * i + k] - xw[3 * i + k];
    std::vector<std::pair<int, std::array<int, 3>>> cand;
    for (int i = 0; i < n; ++i) {
        const int bi = bin_of[i];
        const int b0 = bi / (nb[1] * nb[2]), b1 = (bi / nb[2]) % nb[1], b2 = bi % nb[2];
        const double* xi = &xw[3 * i];
        for (int d0 = -reach[0]; d0 <= reach[0]; ++d0)
            for (int d1 = -reach[1]; d1 <= reach[1]; ++d1)
                for (int d2 = -reach[2]; d2 <= reach[2]; ++d2) {
                    const int c0 = b0 + d0, c1 = b1 + d1, c2 = b2 + d2;
                    // bin index and the lattice image it lies in
                    const int i0 = ((c0 % nb[0]) + nb[0]) % nb[0], s0 = (c0 - i0) / nb[0];
                    const int i1 = ((c1 % nb[1]) + nb[1]) % nb[1], s1 = (c1 - i1) / nb[1];
                    const int i2 = ((c2 % nb[2]) + nb[2]) % nb[2], s2 = (c2 - i2) / nb[2];
                    const int bj = (i0 * nb[1] + i1) * nb[2] + i2;
                    double sv[3];
                    for (int k = 0; k < 3; ++k) sv[k] = s0 * cell[k] + s1 * cell[3 + k] + s2 * cell[6 + k];
                    for (int q = head[bj]; q < head[bj + 1]; ++q) {
                        const int j = members[q];
                        // neighbor at x_j + S: del = x_i - x_j - S
                        const double dx = xi[0] - xw[3 * j] - sv[0];
                        const double dy = xi[1] - xw[3 * j + 1] - sv[1];
                        const double dz = xi[2] - xw[3 * j + 2] - sv[2];
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 < rc2 && r2 > 0.0) {
                            out.idx.push_back(j);
                            // del = x_i - x_j + shift on the caller's (unwrapped) positions
                            const double* wi = &out.wrap_shift[3 * i];
                            const double* wj = &out.wrap_shift[3 * j];
                            out.shift.push_back(-sv[0] - wi[0] + wj[0]);
                            out.shift.push_back(-sv[1] - wi[1] + wj[1]);
                            out.shift.push_back(-sv[2] - wi[2] + wj[2]);
                        }
                    }
                }
        out.off[i + 1] = static_cast<int64_t>(out.idx.size());
    }
    return true;
}

}  // namespace mlip
