// TEST INFRASTRUCTURE ONLY -- the CPU oracle of the radial-only style.  Only
// tests/ may load it; the product path never links or calls it.
//
// FP64 C++/OpenMP restatement of the reference CPU pair style
// `pair_style mlip_pair` (src/USER-MLIP/pair_mlip_pair.cpp) with the model
// classes it uses, keeping the reference's loop structure (per-atom
// features accumulated with `omp atomic` over a half list, unique partial
// products, per-pair energy and force, ev_tally accumulation):
//   DataMLIPBase::initialize (des_type pair)  mlip_model.h:35-120       -> PairPot::load
//   ModelParams ctor, combination2/3          mlip_model_params.cpp:27-82,172-221
//   check_type_comb_pair                      mlip_model_params.cpp:212-221
//   Polynomial ctor (model_type 1 combs)      mlip_polynomial.cpp:27-55
//   PolynomialPair ctor, set_uniq_comb,
//   set_polynomial_array                      mlip_polynomial_pair.cpp:29-92
//   type_comb_ table                          mlip_model.h:106-118
//   get_fn                                    mlip_features.cpp:26-52
//   PairMLIPPair::compute and callees         pair_mlip_pair.cpp:56-258
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {

using vec1i = std::vector<int>;
using vec2i = std::vector<vec1i>;
using vec1d = std::vector<double>;
using vec2d = std::vector<vec1d>;

constexpr double kPi = 3.141592653589793;  // mlip_basis_function.cpp:26

bool next_line(std::istream& in, std::string& line) {
    while (std::getline(in, line)) {
        if (!line.empty() && line.find_first_not_of(" \t\r") != std::string::npos) return true;
    }
    return false;
}

template <typename T>
bool first(std::istream& in, T& v) {
    std::string line;
    if (!next_line(in, line)) return false;
    std::istringstream ss(line);
    return static_cast<bool>(ss >> v);
}

template <typename T>
bool all_of_line(std::istream& in, std::vector<T>& out, size_t n) {
    std::string line;
    if (!next_line(in, line)) return false;
    std::istringstream ss(line);
    out.clear();
    T v;
    while (out.size() < n && ss >> v) out.push_back(v);
    return out.size() == n;
}

struct PolyTerm {
    int n0, reg_i, comb_i, order;
};

struct PairPot {
    std::vector<std::string> elements;
    double cutoff = 0.0;
    std::string pair_type, des_type;
    int model_type = 0, maxp = 0, maxl = 0;
    vec1d reg;           // 2 * coeff / scale (mlip_model.h:90)
    vec1d beta, mu;
    vec1d masses;
    int n_type = 0, n_fn = 0, n_des = 0, n_coeff = 0;
    std::vector<vec2i> type_comb_pair;  // [tc][type1] -> partner types
    vec2i type_comb;                    // [type1][type2] -> tc
    vec2i comb2, comb3;
    std::vector<std::vector<std::vector<PolyTerm>>> poly;  // [tc][n]
    vec2i uniq_comb;

    bool check_tc(const vec1i& idx, int type1) const {
        size_t mn = SIZE_MAX;
        for (int p : idx) mn = std::min(mn, type_comb_pair[p][type1].size());
        return mn > 0;
    }

    std::string load(const std::string& path) {
        std::ifstream in(path);
        if (!in) return "cannot open " + path;
        std::string line;
        if (!next_line(in, line)) return "empty file";
        {
            std::istringstream ss(line);
            std::vector<std::string> toks;
            std::string t;
            while (ss >> t) toks.push_back(t);
            if (toks.size() < 3) return "bad element line";
            elements.assign(toks.begin(), toks.end() - 2);
        }
        n_type = static_cast<int>(elements.size());
        if (!first(in, cutoff) || !first(in, pair_type) || !first(in, des_type) || !first(in, model_type) ||
            !first(in, maxp) || !first(in, maxl))
            return "truncated header";
        if (des_type != "pair") return "des_type must be pair";
        if (!first(in, n_coeff)) return "missing n_coeff";
        vec1d coeffs, scales;
        if (!all_of_line(in, coeffs, n_coeff) || !all_of_line(in, scales, n_coeff)) return "bad coefficient lines";
        int n_params = 0;
        if (!first(in, n_params)) return "missing n_params";
        for (int k = 0; k < n_params; ++k) {
            vec1d bm;
            if (!all_of_line(in, bm, 2)) return "bad pair params";
            beta.push_back(bm[0]);
            mu.push_back(bm[1]);
        }
        if (!all_of_line(in, masses, n_type)) return "bad masses";
        reg.resize(n_coeff);
        for (int k = 0; k < n_coeff; ++k) reg[k] = 2.0 * coeffs[k] / scales[k];
        n_fn = n_params;
        // ModelParams (mlip_model_params.cpp:27-82)
        if (n_type == 1)
            type_comb_pair = {{{0}}};
        else if (n_type == 2)
            type_comb_pair = {{{0}, {}}, {{1}, {0}}, {{}, {1}}};
        else
            return "n_type must be 1 or 2";
        const int n_tc = static_cast<int>(type_comb_pair.size());
        n_des = n_fn * n_tc;
        vec1i pindex;
        if (model_type == 2)
            for (int k = 0; k < n_des; ++k) pindex.push_back(k);
        if (model_type == 1) {
            for (int k = 0; k < n_des; ++k) {
                if (maxp > 1) comb2.push_back({k, k});
                if (maxp > 2) comb3.push_back({k, k, k});
            }
        } else {
            auto tcof = [&](int seq) { return seq / n_fn; };
            if (maxp > 1)
                for (size_t i1 = 0; i1 < pindex.size(); ++i1)
                    for (size_t i2 = 0; i2 <= i1; ++i2)
                        for (int t1 = 0; t1 < n_type; ++t1)
                            if (check_tc({tcof(pindex[i1]), tcof(pindex[i2])}, t1)) {
                                comb2.push_back({pindex[i2], pindex[i1]});
                                break;
                            }
            if (maxp > 2)
                for (size_t i1 = 0; i1 < pindex.size(); ++i1)
                    for (size_t i2 = 0; i2 <= i1; ++i2)
                        for (size_t i3 = 0; i3 <= i2; ++i3)
                            for (int t1 = 0; t1 < n_type; ++t1)
                                if (check_tc({tcof(pindex[i1]), tcof(pindex[i2]), tcof(pindex[i3])}, t1)) {
                                    comb3.push_back({pindex[i3], pindex[i2], pindex[i1]});
                                    break;
                                }
        }
        const int expect = model_type == 1 ? n_des * maxp : n_des + static_cast<int>(comb2.size() + comb3.size());
        if (expect != n_coeff) return "coefficient count does not match the model";
        // PolynomialPair (mlip_polynomial_pair.cpp:29-92)
        poly.assign(n_tc, std::vector<std::vector<PolyTerm>>(n_fn));
        std::set<vec1i> uniq;
        auto perms = [](int order) {
            vec2i out(order);
            for (int i = 0; i < order; ++i) {
                out[i].push_back(i);
                for (int j = 0; j < order; ++j)
                    if (j != i) out[i].push_back(j);
            }
            return out;
        };
        for (const vec2i* cs : {&comb2, &comb3})
            for (const auto& c : *cs)
                for (const auto& seq : perms(static_cast<int>(c.size()))) {
                    vec1i o;
                    for (size_t i = 1; i < c.size(); ++i) o.push_back(c[seq[i]]);
                    uniq.insert(o);
                }
        int reg_i = n_des;
        for (const vec2i* cs : {&comb2, &comb3})
            for (const auto& c : *cs) {
                const int order = static_cast<int>(c.size());
                for (const auto& seq : perms(order)) {
                    vec1i o;
                    for (size_t i = 1; i < c.size(); ++i) o.push_back(c[seq[i]]);
                    std::sort(o.begin(), o.end());
                    const int ci = static_cast<int>(std::distance(uniq.begin(), uniq.find(o)));
                    const int tc0 = c[seq[0]] / n_fn, n0 = c[seq[0]] % n_fn;
                    poly[tc0][n0].push_back({n0, reg_i, ci, order});
                }
                ++reg_i;
            }
        uniq_comb.assign(uniq.begin(), uniq.end());
        // type_comb_ (mlip_model.h:106-118)
        type_comb.assign(n_type, vec1i(n_type, 0));
        for (int t1 = 0; t1 < n_type; ++t1)
            for (int t2 = 0; t2 < n_type; ++t2)
                for (int i = 0; i < n_tc; ++i)
                    if (!type_comb_pair[i][t1].empty() && type_comb_pair[i][t1][0] == t2) {
                        type_comb[t1][t2] = i;
                        break;
                    }
        return "";
    }

    // mlip_features.cpp:26-52
    void fn(double r, vec1d& g, vec1d* gd) const {
        const double fc = 0.5 * (std::cos(kPi * r / cutoff) + 1.0);
        const double fcd = -0.5 * kPi / cutoff * std::sin(kPi * r / cutoff);
        g.resize(n_fn);
        if (gd) gd->resize(n_fn);
        for (int n = 0; n < n_fn; ++n) {
            const double d = r - mu[n];
            const double e = std::exp(-beta[n] * d * d);
            g[n] = e * fc;
            if (gd) (*gd)[n] = -2.0 * beta[n] * d * e * fc + e * fcd;
        }
    }
};

struct PairOracle {
    PairPot pot;
    vec1i types_by_tag;
};

}  // namespace

extern "C" {

void* pair_oracle_create(const char* potfile, char* err, int errlen) {
    auto* o = new PairOracle();
    const std::string e = o->pot.load(potfile);
    if (!e.empty()) {
        if (err && errlen > 0) std::snprintf(err, errlen, "%s", e.c_str());
        delete o;
        return nullptr;
    }
    return o;
}

void pair_oracle_destroy(void* h) { delete static_cast<PairOracle*>(h); }

// n_type, n_fn, n_tc, n_des, n_coeff, model_type, maxp, n_comb2, n_comb3
int pair_oracle_info(void* h, int* out) {
    const PairPot& p = static_cast<PairOracle*>(h)->pot;
    const int v[9] = {p.n_type, p.n_fn, static_cast<int>(p.type_comb_pair.size()), p.n_des, p.n_coeff,
                      p.model_type, p.maxp, static_cast<int>(p.comb2.size()), static_cast<int>(p.comb3.size())};
    std::memcpy(out, v, sizeof v);
    return 9;
}

double pair_oracle_cutoff(void* h) { return static_cast<PairOracle*>(h)->pot.cutoff; }

int pair_oracle_masses(void* h, double* out) {
    const PairPot& p = static_cast<PairOracle*>(h)->pot;
    for (size_t k = 0; k < p.masses.size(); ++k) out[k] = p.masses[k];
    return static_cast<int>(p.masses.size());
}

int pair_oracle_combs(void* h, int order, int* out) {
    const PairPot& p = static_cast<PairOracle*>(h)->pot;
    const vec2i& c = order == 2 ? p.comb2 : p.comb3;
    int k = 0;
    for (const auto& t : c)
        for (int v : t) out[k++] = v;
    return static_cast<int>(c.size());
}

void pair_oracle_set_types_by_tag(void* h, int n, const int* elem_of_tag) {
    static_cast<PairOracle*>(h)->types_by_tag.assign(elem_of_tag, elem_of_tag + n);
}

// PairMLIPPair::compute (pair_mlip_pair.cpp:56-258) on ilist/numneigh plus a
// flattened firstneigh.  f (nall*3) accumulated; eng and virial overwritten;
// feat_out (nlocal x n_des, nullable) receives the per-atom features.
int pair_oracle_compute(void* h, int nlocal, int nall, const double* x, const int64_t* tag, int inum,
                        const int* ilist, const int* numneigh, const int* neigh_off, const int* neigh_idx,
                        double* f, double* eng, double* virial, double* eatom, double* vatom, double* feat_out) {
    auto* o = static_cast<PairOracle*>(h);
    const PairPot& pot = o->pot;
    if (static_cast<int>(o->types_by_tag.size()) < nlocal) return -1;
    (void)nall;
    const vec1i& types = o->types_by_tag;
    const int n_fn = pot.n_fn, n_tc = static_cast<int>(pot.type_comb_pair.size());
    auto del = [&](int i, int j, double& dx, double& dy, double& dz) {
        dx = x[3 * i] - x[3 * j];
        dy = x[3 * i + 1] - x[3 * j + 1];
        dz = x[3 * i + 2] - x[3 * j + 2];
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    };
    // first part of the polynomial correction: features, unique products
    std::vector<vec2d> pf(n_tc, vec2d(nlocal)), pe(n_tc, vec2d(nlocal));
    vec2d dn(nlocal, vec1d(pot.n_des, 0.0));
    if (pot.maxp > 1 || feat_out) {
#pragma omp parallel for schedule(guided)
        for (int ii = 0; ii < inum; ++ii) {
            const int i = ilist[ii];
            const int t1 = types[tag[i] - 1];
            vec1d g;
            for (int jj = 0; jj < numneigh[i]; ++jj) {
                const int j = neigh_idx[neigh_off[ii] + jj];
                double dx, dy, dz;
                const double r = del(i, j, dx, dy, dz);
                if (r < pot.cutoff) {
                    const int t2 = types[tag[j] - 1];
                    const int s = pot.type_comb[t1][t2] * n_fn;
                    pot.fn(r, g, nullptr);
                    for (int n = 0; n < n_fn; ++n) {
#pragma omp atomic
                        dn[tag[i] - 1][s + n] += g[n];
#pragma omp atomic
                        dn[tag[j] - 1][s + n] += g[n];
                    }
                }
            }
        }
    }
    if (pot.maxp > 1) {
#pragma omp parallel for schedule(guided)
        for (int ii = 0; ii < inum; ++ii) {
            const int i = ilist[ii];
            const int t1 = types[tag[i] - 1];
            const vec1d& d = dn[tag[i] - 1];
            vec1d prod(pot.uniq_comb.size(), 0.5);
            for (size_t k = 0; k < pot.uniq_comb.size(); ++k)
                for (int c : pot.uniq_comb[k]) prod[k] *= d[c];
            for (int t2 = 0; t2 < pot.n_type; ++t2) {
                const int tc = pot.type_comb[t1][t2];
                vec1d vf(n_fn, 0.0), ve(n_fn, 0.0);
                for (int n = 0; n < n_fn; ++n)
                    for (const PolyTerm& pt : pot.poly[tc][n]) {
                        const double v = prod[pt.comb_i] * pot.reg[pt.reg_i];
                        vf[n] += v;
                        ve[n] += v / pt.order;
                    }
                pf[tc][tag[i] - 1] = vf;
                pe[tc][tag[i] - 1] = ve;
            }
        }
    }
    // per-pair energy and force, then ev_tally in list order
    vec2d ev(inum), fp(inum);
#pragma omp parallel for schedule(guided)
    for (int ii = 0; ii < inum; ++ii) {
        const int i = ilist[ii];
        const int t1 = types[tag[i] - 1];
        ev[ii].assign(numneigh[i], 0.0);
        fp[ii].assign(numneigh[i], 0.0);
        vec1d g, gd;
        for (int jj = 0; jj < numneigh[i]; ++jj) {
            const int j = neigh_idx[neigh_off[ii] + jj];
            double dx, dy, dz;
            const double r = del(i, j, dx, dy, dz);
            if (r < pot.cutoff) {
                const int t2 = types[tag[j] - 1];
                const int tc = pot.type_comb[t1][t2], s = tc * n_fn;
                pot.fn(r, g, &gd);
                double fpair = 0.0, e = 0.0;
                for (int n = 0; n < n_fn; ++n) {
                    fpair += gd[n] * pot.reg[s + n];
                    e += g[n] * pot.reg[s + n];
                }
                if (pot.maxp > 1) {
                    for (int n = 0; n < n_fn; ++n) {
                        fpair += gd[n] * pf[tc][tag[i] - 1][n] + gd[n] * pf[tc][tag[j] - 1][n];
                        e += g[n] * pe[tc][tag[i] - 1][n] + g[n] * pe[tc][tag[j] - 1][n];
                    }
                }
                fpair *= -1.0 / r;
                ev[ii][jj] = e;
                fp[ii][jj] = fpair;
            }
        }
    }
    double E = 0.0, V[6] = {0, 0, 0, 0, 0, 0};
    for (int ii = 0; ii < inum; ++ii) {
        const int i = ilist[ii];
        for (int jj = 0; jj < numneigh[i]; ++jj) {
            const int j = neigh_idx[neigh_off[ii] + jj];
            double dx, dy, dz;
            const double r = del(i, j, dx, dy, dz);
            if (r < pot.cutoff) {
                const double fpair = fp[ii][jj];
                f[3 * i] += fpair * dx;
                f[3 * i + 1] += fpair * dy;
                f[3 * i + 2] += fpair * dz;
                f[3 * j] -= fpair * dx;
                f[3 * j + 1] -= fpair * dy;
                f[3 * j + 2] -= fpair * dz;
                E += ev[ii][jj];
                const double v[6] = {dx * dx * fpair, dy * dy * fpair, dz * dz * fpair,
                                     dx * dy * fpair, dx * dz * fpair, dy * dz * fpair};
                for (int k = 0; k < 6; ++k) V[k] += v[k];
                // ev_tally with newton_pair: half of the pair's energy and
                // virial to each atom
                if (eatom) {
                    eatom[i] += 0.5 * ev[ii][jj];
                    eatom[j] += 0.5 * ev[ii][jj];
                }
                if (vatom)
                    for (int k = 0; k < 6; ++k) {
                        vatom[6 * static_cast<size_t>(i) + k] += 0.5 * v[k];
                        vatom[6 * static_cast<size_t>(j) + k] += 0.5 * v[k];
                    }
            }
        }
    }
    *eng = E;
    for (int k = 0; k < 6; ++k) virial[k] = V[k];
    if (feat_out)
        for (int i = 0; i < nlocal; ++i)
            for (int k = 0; k < pot.n_des; ++k) feat_out[static_cast<size_t>(i) * pot.n_des + k] = dn[i][k];
    return 0;
}

}  // extern "C"
