// Synthetic structures for the batched and served workloads: random
// triclinic cells at a given number density with uniformly random positions
// under a minimum-distance rejection (periodic images included), two species
// with a per-structure random fraction.  Deterministic from (seed, index)
// through a splitmix64-seeded xoshiro256** stream.
#include "workload.h"

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace mlip {

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    Rng(uint64_t seed, uint64_t stream) {
        uint64_t x = seed * 0x9E3779B97F4A7C15ull ^ (stream + 1) * 0xD1B54A32D192ED03ull;
        for (auto& v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

int random_structure_size(uint64_t seed, int index, int nmin, int nmax) {
    Rng r(seed, static_cast<uint64_t>(index) * 2);
    const double lo = std::log(static_cast<double>(nmin)), hi = std::log(static_cast<double>(nmax));
    int n = static_cast<int>(std::lround(std::exp(lo + (hi - lo) * r.uniform())));
    return std::min(nmax, std::max(nmin, n));
}

bool random_structure(uint64_t seed, int index, int natoms, double density, double rmin, double cell[9],
                      double* x, int* type) {
    Rng r(seed, static_cast<uint64_t>(index) * 2 + 1);
    const double vol = natoms / density;
    // random triclinic cell: edge ratios in [0.8, 1.25], angles in [70, 110] deg
    for (int tries = 0;; ++tries) {
        double len[3], ang[3];
        for (int k = 0; k < 3; ++k) len[k] = 0.8 + 0.45 * r.uniform();
        for (int k = 0; k < 3; ++k) ang[k] = (70.0 + 40.0 * r.uniform()) * 3.14159265358979323846 / 180.0;
        const double ca = std::cos(ang[0]), cb = std::cos(ang[1]), cg = std::cos(ang[2]), sg = std::sin(ang[2]);
        const double vu2 = 1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg;
        if (vu2 < 0.25 && tries < 100) continue;
        const double vu = std::sqrt(std::max(vu2, 1e-6));
        const double sc = std::cbrt(vol / (len[0] * len[1] * len[2] * vu));
        const double la = len[0] * sc, lb = len[1] * sc, lc = len[2] * sc;
        const double cx = lc * cb, cy = lc * (ca - cb * cg) / sg;
        const double cz = std::sqrt(std::max(lc * lc - cx * cx - cy * cy, 1e-12));
        const double c9[9] = {la, 0, 0, lb * cg, lb * sg, 0, cx, cy, cz};
        std::copy(c9, c9 + 9, cell);
        break;
    }
    // bins in fractional space at least rmin wide along each face normal
    const double* a = cell;
    const double* b = cell + 3;
    const double* c = cell + 6;
    auto cross = [](const double* u, const double* v, double* o) {
        o[0] = u[1] * v[2] - u[2] * v[1];
        o[1] = u[2] * v[0] - u[0] * v[2];
        o[2] = u[0] * v[1] - u[1] * v[0];
    };
    double t[3];
    double w[3];
    cross(b, c, t);
    w[0] = vol / std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    cross(c, a, t);
    w[1] = vol / std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    cross(a, b, t);
    w[2] = vol / std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    int nb[3];
    for (int k = 0; k < 3; ++k) nb[k] = std::max(1, static_cast<int>(std::floor(w[k] / rmin)));
    int reach[3];
    for (int k = 0; k < 3; ++k) reach[k] = static_cast<int>(std::ceil(rmin / (w[k] / nb[k])));
    std::vector<std::vector<int>> bins(static_cast<size_t>(nb[0]) * nb[1] * nb[2]);
    std::vector<double> frac(3 * static_cast<size_t>(natoms));
    const double r2 = rmin * rmin;
    long attempts = 0;
    for (int i = 0; i < natoms;) {
        if (++attempts > 2000000L) return false;
        double f[3] = {r.uniform(), r.uniform(), r.uniform()};
        int bi[3];
        for (int k = 0; k < 3; ++k) bi[k] = std::min(nb[k] - 1, static_cast<int>(f[k] * nb[k]));
        bool ok = true;
        for (int d0 = -reach[0]; d0 <= reach[0] && ok; ++d0)
            for (int d1 = -reach[1]; d1 <= reach[1] && ok; ++d1)
                for (int d2 = -reach[2]; d2 <= reach[2] && ok; ++d2) {
                    const int q0 = bi[0] + d0, q1 = bi[1] + d1, q2 = bi[2] + d2;
                    const int i0 = ((q0 % nb[0]) + nb[0]) % nb[0], i1 = ((q1 % nb[1]) + nb[1]) % nb[1],
                              i2 = ((q2 % nb[2]) + nb[2]) % nb[2];
                    const double s0 = (q0 - i0) / nb[0], s1 = (q1 - i1) / nb[1], s2 = (q2 - i2) / nb[2];
                    for (int j : bins[(static_cast<size_t>(i0) * nb[1] + i1) * nb[2] + i2]) {
                        const double g0 = f[0] - frac[3 * j] - s0, g1 = f[1] - frac[3 * j + 1] - s1,
                                     g2 = f[2] - frac[3 * j + 2] - s2;
                        const double dx = g0 * a[0] + g1 * b[0] + g2 * c[0];
                        const double dy = g0 * a[1] + g1 * b[1] + g2 * c[1];
                        const double dz = g0 * a[2] + g1 * b[2] + g2 * c[2];
                        if (dx * dx + dy * dy + dz * dz < r2) {
                            ok = false;
                            break;
                        }
                    }
                }
        if (!ok) continue;
        for (int k = 0; k < 3; ++k) frac[3 * i + k] = f[k];
        bins[(static_cast<size_t>(bi[0]) * nb[1] + bi[1]) * nb[2] + bi[2]].push_back(i);
        ++i;
    }
    for (int i = 0; i < natoms; ++i)
        for (int k = 0; k < 3; ++k)
            x[3 * i + k] = frac[3 * i] * a[k] + frac[3 * i + 1] * b[k] + frac[3 * i + 2] * c[k];
    // species: a random fraction of the atoms is type 2
    const double fr = r.uniform();
    const int n2 = static_cast<int>(std::lround(fr * natoms));
    std::vector<int> perm(natoms);
    for (int i = 0; i < natoms; ++i) perm[i] = i;
    for (int i = natoms - 1; i > 0; --i) std::swap(perm[i], perm[static_cast<int>(r.next() % (i + 1))]);
    for (int i = 0; i < natoms; ++i) type[i] = 1;
    for (int k = 0; k < n2; ++k) type[perm[k]] = 2;
    return true;
}

}  // namespace mlip
