// Host-side model layer: parser, index enumeration and device tables.
// See model.h for the reference functions each part corresponds to.
#include "model.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <tuple>

extern "C" {
extern const char mlip_gtinv_table_begin[];
extern const char mlip_gtinv_table_end[];
}

namespace mlip {

#include "q34_signatures.inc"
static_assert(MLIP_CAP3 == kUnrollCap3 && MLIP_CAP4 == kUnrollCap4,
              "unroll caps of the model tables and of the order-3/4 kernel's signature list");

namespace {

// (order, output L, l_a, l_b[, l_c]) is one of the signatures the order-3/4
// kernel has unrolled code for (the list q34_signatures.inc compiles into it)
bool unrolled_signature(int q, int L, int la, int lb, int lc) {
    static const std::set<std::array<int, 5>> sigs = [] {
        std::set<std::array<int, 5>> s;
#define MLIP_HOST_SIG3(A, B) s.insert({3, LCUR, A, B, 0});
#define MLIP_HOST_SIG4(A, B, C) s.insert({4, LCUR, A, B, C});
        int LCUR = 0;
        LCUR = 0; MLIP_SIG3_L0(MLIP_HOST_SIG3) MLIP_SIG4_L0(MLIP_HOST_SIG4)
        LCUR = 1; MLIP_SIG3_L1(MLIP_HOST_SIG3) MLIP_SIG4_L1(MLIP_HOST_SIG4)
        LCUR = 2; MLIP_SIG3_L2(MLIP_HOST_SIG3) MLIP_SIG4_L2(MLIP_HOST_SIG4)
        LCUR = 3; MLIP_SIG3_L3(MLIP_HOST_SIG3)
        LCUR = 4; MLIP_SIG3_L4(MLIP_HOST_SIG3)
        LCUR = 5; MLIP_SIG3_L5(MLIP_HOST_SIG3)
        LCUR = 6; MLIP_SIG3_L6(MLIP_HOST_SIG3)
        LCUR = 7; MLIP_SIG3_L7(MLIP_HOST_SIG3)
        LCUR = 8; MLIP_SIG3_L8(MLIP_HOST_SIG3)
#undef MLIP_HOST_SIG3
#undef MLIP_HOST_SIG4
        return s;
    }();
    return sigs.count({q, L, la, lb, q == 4 ? lc : 0}) > 0;
}

}  // namespace

// ------------------------------------------------------------ gtinv table
bool parse_gtinv_text(const char* text, size_t len, GtinvTable& out, std::string& err) {
    std::istringstream in(std::string(text, len));
    std::string tag;
    int n = 0;
    if (!(in >> tag >> n) || tag != "gtinv_table" || n <= 0) {
        err = "gtinv table: bad header";
        return false;
    }
    out.irreps.assign(n, Irrep());
    for (int i = 0; i < n; ++i) {
        int idx = -1, q = 0, nt = 0;
        if (!(in >> tag >> idx >> q) || tag != "irrep" || idx != i || q < 1 || q > 6) {
            err = "gtinv table: bad irrep header";
            return false;
        }
        Irrep& r = out.irreps[i];
        r.l.resize(q);
        for (int k = 0; k < q; ++k) in >> r.l[k];
        in >> nt;
        r.m.assign(nt, std::vector<int>(q));
        r.c.resize(nt);
        for (int t = 0; t < nt; ++t) {
            for (int k = 0; k < q; ++k) in >> r.m[t][k];
            std::string cs;
            in >> cs;
            r.c[t] = std::strtod(cs.c_str(), nullptr);
        }
        if (!in) {
            err = "gtinv table: truncated";
            return false;
        }
    }
    return true;
}

bool load_builtin_gtinv(GtinvTable& out, std::string& err) {
    const size_t len = static_cast<size_t>(mlip_gtinv_table_end - mlip_gtinv_table_begin);
    return parse_gtinv_text(mlip_gtinv_table_begin, len, out, err);
}

// ------------------------------------------------------------ file parsing
namespace {

bool next_line(std::istream& in, std::string& line) { return static_cast<bool>(std::getline(in, line)); }

template <typename T>
bool first_token(std::istream& in, T& v) {
    std::string line;
    if (!next_line(in, line)) return false;
    std::istringstream ss(line);
    return static_cast<bool>(ss >> v);
}

template <typename T>
bool n_tokens(std::istream& in, std::vector<T>& v, int n) {
    std::string line;
    if (!next_line(in, line)) return false;
    std::istringstream ss(line);
    v.resize(n);
    for (int i = 0; i < n; ++i)
        if (!(ss >> v[i])) return false;
    return true;
}

// tc-tuple validity: some centre type sees every type combination
bool tuple_valid(const Model& m, const std::vector<int>& tcs) {
    for (int t = 0; t < m.n_type; ++t) {
        bool ok = true;
        for (int tc : tcs)
            if (m.tcp_other[tc][t] < 0) {
                ok = false;
                break;
            }
        if (ok) return true;
    }
    return false;
}

unsigned mask_of(const Model& m, const std::vector<int>& tcs) {
    unsigned mask = 0;
    for (int t = 0; t < m.n_type; ++t) {
        bool ok = true;
        for (int tc : tcs)
            if (m.tcp_other[tc][t] < 0) ok = false;
        if (ok) mask |= 1u << t;
    }
    return mask;
}

// Features of each irrep: unique multisets {(l_k, tc_k)} over valid tc
// tuples, in ascending lexicographic order of the sorted pair sequence
// (the order std::set<std::multiset<pair>> yields in the reference,
// mlip_model_params.cpp:100-124).
void enumerate_itcs(Model& m) {
    m.itcs.clear();
    for (size_t i = 0; i < m.irreps.size(); ++i) {
        const std::vector<int>& ls = m.irreps[i].l;
        const int q = static_cast<int>(ls.size());
        std::set<std::vector<std::pair<int, int>>> uniq;
        std::vector<int> tup(q, 0);
        while (true) {
            if (tuple_valid(m, tup)) {
                std::vector<std::pair<int, int>> key(q);
                for (int k = 0; k < q; ++k) key[k] = {ls[k], tup[k]};
                std::sort(key.begin(), key.end());
                uniq.insert(key);
            }
            int k = q - 1;
            while (k >= 0 && ++tup[k] == m.n_tc) tup[k--] = 0;
            if (k < 0) break;
        }
        for (const auto& key : uniq) {
            Itc it{};
            it.irrep = static_cast<int>(i);
            it.order = q;
            std::vector<int> tcs(q);
            for (int k = 0; k < q; ++k) {
                tcs[k] = key[k].second;
                it.tc[k] = key[k].second;
            }
            for (int k = q; k < 6; ++k) it.tc[k] = -1;
            it.type_mask = mask_of(m, tcs);
            m.itcs.push_back(it);
        }
    }
    m.n_gtinv = static_cast<int>(m.itcs.size());
}

bool share_type(unsigned a, unsigned b) { return (a & b) != 0; }

void enumerate_polynomial(Model& m) {
    m.n_des = m.n_fn * m.n_gtinv;
    m.comb2.clear();
    m.comb3.clear();
    m.poly_features.clear();
    if (m.model_type == 1) {
        // products of a feature with itself (mlip_polynomial.cpp:36-40)
        for (int f = 0; f < m.n_des; ++f) {
            if (m.maxp > 1) m.comb2.push_back({f, f});
            if (m.maxp > 2) m.comb3.push_back({f, f, f});
        }
        m.n_coeff = m.n_des * m.maxp;
        if (m.maxp > 1)
            for (int f = 0; f < m.n_des; ++f) m.poly_features.push_back(f);
        return;
    }
    std::vector<int> pidx;
    for (int f = 0; f < m.n_des; ++f) {
        const Itc& it = m.itcs[f % m.n_gtinv];
        const bool use = (m.model_type == 2) || (m.model_type == 3 && it.order == 1) ||
                         (m.model_type == 4 && it.order < 3);
        if (use) pidx.push_back(f);
    }
    auto mask = [&](int f) { return m.itcs[f % m.n_gtinv].type_mask; };
    if (m.maxp > 1)
        for (size_t i1 = 0; i1 < pidx.size(); ++i1)
            for (size_t i2 = 0; i2 <= i1; ++i2)
                if (share_type(mask(pidx[i1]), mask(pidx[i2]))) m.comb2.push_back({pidx[i2], pidx[i1]});
    if (m.maxp > 2)
        for (size_t i1 = 0; i1 < pidx.size(); ++i1)
            for (size_t i2 = 0; i2 <= i1; ++i2)
                for (size_t i3 = 0; i3 <= i2; ++i3)
                    if (mask(pidx[i1]) & mask(pidx[i2]) & mask(pidx[i3]))
                        m.comb3.push_back({pidx[i3], pidx[i2], pidx[i1]});
    m.n_coeff = m.n_des + static_cast<int>(m.comb2.size() + m.comb3.size());
    if (m.maxp > 1) m.poly_features = pidx;
}

double lookup(const std::map<std::vector<int>, double>& c, const std::vector<int>& mt) {
    auto it = c.find(mt);
    return it == c.end() ? 0.0 : it->second;
}

// Species pairing of the adjoint rows (two-species models): for L <=
// min(kPairMaxL, plan_pair), rows (0, L) and (1, L) whose visits all run unrolled become one
// row (S = kPairedRows).  A visit of row 0 and one of row 1 with the same
// order, factors (l and species), energy class and coefficient block (the
// same irrep seen from two families that differ only in the output species)
// become one visit of both rows: the kernel forms its products once and
// adds them with each family's H.  The rest keep their row.  Within the
// paired row the visits go by energy class, and inside a class row-0 visits,
// row-1 visits, then shared ones, so each row still sees its classes in
// order.  `pend` arrives sorted by (S, L, class, ...).
void pair_rows(const Model& m, std::vector<PlanVisit>& pend) {
    if (m.n_type != 2) return;
    auto cls = [](const VisitQ34& v) { return v.pslot >= 0 ? 4 : std::max(v.order - 1, 0); };
    std::vector<PlanVisit> out;
    std::vector<bool> taken(pend.size(), false);
    for (int L = 0; L <= std::min(kPairMaxL, m.plan_pair); ++L) {
        std::vector<size_t> r0, r1;
        bool all_fast = true;
        for (size_t k = 0; k < pend.size(); ++k) {
            if (pend[k].L != L) continue;
            if (!(pend[k].v.flags & 4)) all_fast = false;
            (pend[k].S == 0 ? r0 : r1).push_back(k);
        }
        if (!all_fast || r0.empty() || r1.empty()) continue;
        std::vector<bool> used1(r1.size(), false);
        std::vector<PlanVisit> only0[5], only1[5], both[5];
        for (size_t i0 : r0) {
            const PlanVisit& a = pend[i0];
            const int c = cls(a.v);
            int hit = -1;
            if (a.v.order >= 3)
                for (size_t j = 0; j < r1.size() && hit < 0; ++j) {
                    if (used1[j]) continue;
                    const PlanVisit& b = pend[r1[j]];
                    const VisitQ34 &x = a.v, &y = b.v;
                    if (y.order != x.order || y.la != x.la || y.sa != x.sa || y.lb != x.lb || y.sb != x.sb ||
                        y.lc != x.lc || y.sc != x.sc || cls(y) != c || b.clen != a.clen)
                        continue;
                    bool same = true;
                    for (int q = 0; q < a.clen && same; ++q) same = m.coef[x.coef + q] == m.coef[y.coef + q];
                    if (same) hit = static_cast<int>(j);
                }
            PlanVisit q = a;
            q.S = kPairedRows;
            if (hit >= 0) {
                used1[hit] = true;
                q.v.row = 2;
                q.v.pslot1 = pend[r1[hit]].v.pslot;
                q.v.itc1 = pend[r1[hit]].v.itc;
                both[c].push_back(q);
            } else {
                q.v.row = 0;
                only0[c].push_back(q);
            }
        }
        for (size_t j = 0; j < r1.size(); ++j) {
            if (used1[j]) continue;
            PlanVisit q = pend[r1[j]];
            q.S = kPairedRows;
            q.v.row = 1;
            only1[cls(q.v)].push_back(q);
        }
        for (size_t k : r0) taken[k] = true;
        for (size_t k : r1) taken[k] = true;
        for (int c = 0; c < 5; ++c) {
            out.insert(out.end(), only0[c].begin(), only0[c].end());
            out.insert(out.end(), only1[c].begin(), only1[c].end());
            out.insert(out.end(), both[c].begin(), both[c].end());
        }
    }
    if (out.empty()) return;
    // unpaired rows first in their (S, L) order, then the paired rows by L
    std::vector<PlanVisit> rest;
    for (size_t k = 0; k < pend.size(); ++k)
        if (!taken[k]) rest.push_back(pend[k]);
    rest.insert(rest.end(), out.begin(), out.end());
    pend.swap(rest);
}

// Group visits by (S, L) and balance the groups over the waves of a
// workgroup.  The work model of a visit is its terms plus region_cost per
// m_a (per (m_a, m_b)) region and visit_cost per visit, in units of terms.
void make_groups(const Model& m, std::vector<PlanVisit> pend, std::vector<GroupQ34>& groups,
                 std::vector<VisitQ34>& visits, bool adjoint, int& red_total, bool& dyn, int region_cost,
                 int visit_cost) {
    // group by (S, L); inside a group, visits by energy class (linear
    // order 1, 2, 3, 4, polynomial: the adjoint kernel takes each
    // linear class's energy from the accumulated G row at the class
    // boundary), then with the same code path (signature) adjacent for
    // instruction-cache locality
    auto cls = [](const VisitQ34& v) { return v.pslot >= 0 ? 4 : std::max(v.order - 1, 0); };
    std::stable_sort(pend.begin(), pend.end(), [&](const PlanVisit& a, const PlanVisit& b) {
        if (a.S != b.S) return a.S < b.S;
        if (a.L != b.L) return a.L < b.L;
        const auto ka = std::make_tuple(cls(a.v), a.v.flags & 4, a.v.order, a.v.la, a.v.lb, a.v.lc);
        const auto kb = std::make_tuple(cls(b.v), b.v.flags & 4, b.v.order, b.v.la, b.v.lb, b.v.lc);
        return ka < kb;
    });
    for (const PlanVisit& q : pend)
        if (!(q.v.flags & 4)) dyn = true;
    if (adjoint) pair_rows(m, pend);
    // Work model of a visit: its terms plus the per-m_a (per (m_a, m_b))
    // region and per-visit overheads, in units of terms (3 and 24 were the
    // best of a sweep at the 256k-atom benchmark model: 40.3 -> 38.8 ms);
    // of a part: plus loading and storing its G row.
    auto vcost = [&](const PlanVisit& q) {
        const long regions = q.v.order <= 2 ? 0 : (2L * q.v.la + 1) * (q.v.order == 4 ? 2L * q.v.lb + 1 : 1);
        // a paired row's visit adds its products into one or both rows:
        // 2 (L + 1) FMAs each, about a third of a term's work per entry
        const long fin = (q.S == kPairedRows && q.v.order >= 3) ? (q.v.row == 2 ? 2L : 1L) * (q.L + 1) / 3 : 0L;
        return q.cost + region_cost * regions + visit_cost + fin;
    };
    struct Row {
        size_t b, e;
        int L, S;
    };
    std::vector<Row> rows;
    for (size_t k = 0; k < pend.size();) {
        size_t e = k;
        while (e < pend.size() && pend[e].S == pend[k].S && pend[e].L == pend[k].L) ++e;
        rows.push_back({k, e, pend[k].L, pend[k].S});
        k = e;
    }
    // LDS available for the partial rows of split rows (adjoint only):
    // what the workgroup's 160 KB leave beside the a_nlm tile
    const long tile = static_cast<long>(m.n_type) * n_lmh_of(m.lmax3) * 2 * kBlockLanes;
    const long budget = adjoint ? std::max(0L, 160L * 1024 / 8 - tile) : 1L << 40;
    // a paired row keeps two partial rows per part
    auto row_red = [](int L, int np, int S) {
        return np > 1 ? static_cast<long>(np) * (L + 1) * 2 * kBlockLanes * (S == kPairedRows ? 2 : 1) : 0L;
    };
    struct Part {
        size_t row, b, e;
        long cost;
    };
    // split row r into np parts balanced by cost at visit boundaries
    auto split = [&](size_t r, int np, std::vector<Part>& out) {
        const Row& R = rows[r];
        long c = 0;
        for (size_t j = R.b; j < R.e; ++j) c += vcost(pend[j]);
        size_t j = R.b;
        long acc = 0;
        for (int part = 0; part < np; ++part) {
            Part P{r, j, j, 3L * (R.L + 1) + 20};
            if (part == np - 1) {
                while (j < R.e) P.cost += vcost(pend[j++]);
            } else {
                const long goal = c * (part + 1) / np;
                const size_t keep = static_cast<size_t>(np - 1 - part);
                do {
                    const long vc = vcost(pend[j++]);
                    P.cost += vc;
                    acc += vc;
                } while (j < R.e && acc < goal && R.e - j > keep);
            }
            P.e = j;
            out.push_back(P);
        }
    };
    // longest-processing-time assignment, then moves and swaps that lower
    // the most loaded wave; returns the wave loads, heaviest first
    auto assign = [&](const std::vector<Part>& parts, std::vector<int>& wave) {
        std::vector<size_t> order(parts.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return parts[a].cost > parts[b].cost; });
        std::vector<long> load(kWavesPerWg, 0);
        wave.assign(parts.size(), 0);
        for (size_t k : order) {
            const int w = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
            wave[k] = w;
            load[w] += parts[k].cost;
        }
        for (int it = 0; it < 200; ++it) {
            const int hi = static_cast<int>(std::max_element(load.begin(), load.end()) - load.begin());
            long best = load[hi];
            int bk = -1, bk2 = -1, bw = -1;
            for (size_t k = 0; k < parts.size(); ++k) {
                if (wave[k] != hi) continue;
                for (int w = 0; w < kWavesPerWg; ++w) {
                    if (w == hi) continue;
                    // move k to w
                    const long mv = std::max(load[hi] - parts[k].cost, load[w] + parts[k].cost);
                    if (mv < best) {
                        best = mv;
                        bk = static_cast<int>(k);
                        bk2 = -1;
                        bw = w;
                    }
                    // swap k with a part k2 of w
                    for (size_t k2 = 0; k2 < parts.size(); ++k2) {
                        if (wave[k2] != w) continue;
                        const long d = parts[k].cost - parts[k2].cost;
                        if (d <= 0) continue;
                        const long sw = std::max(load[hi] - d, load[w] + d);
                        if (sw < best) {
                            best = sw;
                            bk = static_cast<int>(k);
                            bk2 = static_cast<int>(k2);
                            bw = w;
                        }
                    }
                }
            }
            if (bk < 0) break;
            load[hi] -= parts[bk].cost;
            load[bw] += parts[bk].cost;
            wave[bk] = bw;
            if (bk2 >= 0) {
                load[bw] -= parts[bk2].cost;
                load[hi] += parts[bk2].cost;
                wave[bk2] = hi;
            }
        }
        std::sort(load.begin(), load.end(), std::greater<long>());
        return load;
    };
    std::vector<int> nparts(rows.size(), 1);
    auto build = [&](std::vector<Part>& parts) {
        parts.clear();
        for (size_t r = 0; r < rows.size(); ++r) split(r, nparts[r], parts);
    };
    std::vector<Part> parts;
    std::vector<int> wave;
    build(parts);
    std::vector<long> best = assign(parts, wave);
    long red_used = 0;
    // split one row one part further while that lowers the wave loads
    // (heaviest first, compared in order: with two rows at the maximum, the
    // first split leaves the maximum and lowers the count of waves at it)
    // and the partial rows fit in LDS
    for (;;) {
        std::vector<long> cand_best = best;
        size_t cand = rows.size();
        for (size_t r = 0; r < rows.size(); ++r) {
            const Row& R = rows[r];
            if (static_cast<size_t>(nparts[r]) >= R.e - R.b || nparts[r] >= kWavesPerWg) continue;
            const long red_new = red_used - row_red(R.L, nparts[r], R.S) + row_red(R.L, nparts[r] + 1, R.S);
            if (red_new > budget) continue;
            ++nparts[r];
            std::vector<Part> p2;
            std::vector<int> w2;
            build(p2);
            const std::vector<long> mx = assign(p2, w2);
            --nparts[r];
            if (mx < cand_best) {
                cand_best = mx;
                cand = r;
            }
        }
        if (cand == rows.size()) break;
        red_used += row_red(rows[cand].L, nparts[cand] + 1, rows[cand].S) - row_red(rows[cand].L, nparts[cand], rows[cand].S);
        ++nparts[cand];
        build(parts);
        best = assign(parts, wave);
    }
    // emit: parts in row order, partial-row offsets per split row
    int red = 0;
    size_t pi = 0;
    for (size_t r = 0; r < rows.size(); ++r) {
        const int np = nparts[r];
        for (int part = 0; part < np; ++part, ++pi) {
            const Part& P = parts[pi];
            GroupQ34 g{};
            g.S = pend[rows[r].b].S;
            g.L = rows[r].L;
            g.part = part;
            g.nparts = np;
            g.red = (adjoint && np > 1)
                        ? red + part * (rows[r].L + 1) * 2 * kBlockLanes * (g.S == kPairedRows ? 2 : 1)
                        : 0;
            g.vbeg = static_cast<int>(visits.size());
            for (size_t j = P.b; j < P.e; ++j) visits.push_back(pend[j].v);
            g.vend = static_cast<int>(visits.size());
            g.wave = wave[pi];
            groups.push_back(g);
        }
        if (adjoint && np > 1) red += static_cast<int>(row_red(rows[r].L, np, rows[r].S));
    }
    red_total = std::max(red_total, red);
}

// Builds the flattened visit tables for centre type t.
void build_type_tables(Model& m, int t, const std::vector<std::map<std::vector<int>, double>>& cmap) {
    TypeTables tt;
    std::vector<std::vector<VisitQ12>> by_l(m.maxl + 1);
    using Pending = PlanVisit;
    std::vector<Pending> adj, feat;
    auto push_coef = [&](double v) { m.coef.push_back(v); };
    for (int i = 0; i < m.n_gtinv; ++i) {
        const Itc& it = m.itcs[i];
        if (!((it.type_mask >> t) & 1u)) continue;
        const Irrep& ir = m.irreps[it.irrep];
        const std::map<std::vector<int>, double>& cm = cmap[it.irrep];
        const int q = it.order;
        std::vector<int> sp(q);
        for (int k = 0; k < q; ++k) sp[k] = m.tcp_other[it.tc[k]][t];
        const int pslot = m.prank[i];
        if (q <= 2) {
            VisitQ12 v{};
            v.order = q;
            v.itc = i;
            v.s0 = sp[0];
            v.s1 = q == 2 ? sp[1] : sp[0];
            v.pslot = pslot;
            v.coef = static_cast<int>(m.coef.size());
            const int L = ir.l[0];
            if (q == 1) {
                push_coef(lookup(cm, {0}));
            } else {
                for (int m0 = -L; m0 <= L; ++m0) push_coef(lookup(cm, {m0, -m0}));
            }
            by_l[L].push_back(v);
            if (L <= m.q12_fuse || (L > m.lmax3 && L <= m.q12_fuse_hi)) {
                // the same family's adjoint in the order-3/4 kernel, one visit
                // per factor position p: dd/da_{sp[p],L,-k} = C a_{other,L,+k}
                // with C = C(-k, k) (p = 0) or C(k, -k) (p = 1)
                for (int pp = 0; pp < q; ++pp) {
                    VisitQ34 f{};
                    f.order = q;
                    f.itc = i;
                    f.la = q == 2 ? L : 0;
                    f.sa = q == 2 ? sp[1 - pp] : 0;
                    f.pslot = pslot;
                    f.flags = 2 | 4;
                    f.coef = static_cast<int>(m.coef.size());
                    f.dcoef = -1;
                    if (q == 1) {
                        push_coef(lookup(cm, {0}));
                    } else {
                        for (int k = 0; k <= L; ++k) push_coef(lookup(cm, pp == 0 ? std::vector<int>{-k, k}
                                                                                   : std::vector<int>{k, -k}));
                    }
                    const int nc = q == 1 ? 1 : L + 1;
                    adj.push_back({sp[pp], L, f, nc, nc});
                }
            }
            continue;
        }
        if (q >= 5) {
            // orders 5 and 6: one visit per factor position, terms listed
            for (int p = 0; p < q; ++p) {
                VisitHi v{};
                v.S = sp[p];
                v.L = ir.l[p];
                v.order = q;
                v.itc = i;
                v.pslot = pslot;
                int no = 0;
                std::vector<int> oth;
                for (int k = 0; k < q; ++k)
                    if (k != p) {
                        v.ls[no++] = ir.l[k] | (sp[k] << 8);
                        oth.push_back(k);
                    }
                for (int k = no; k < 5; ++k) v.ls[k] = 0;
                auto emit_terms = [&](bool feat) {
                    VisitHi w = v;
                    w.tbeg = static_cast<int32_t>(m.hi_term.size());
                    for (size_t ti = 0; ti < ir.m.size(); ++ti) {
                        const std::vector<int>& mt = ir.m[ti];
                        if (mt[p] > 0) continue;  // outputs m <= 0 only
                        double c = ir.c[ti];
                        if (feat && mt[p] < 0) {
                            std::vector<int> neg(mt);
                            for (int& x : neg) x = -x;
                            c += lookup(cm, neg);
                        }
                        int32_t code = -mt[p];
                        for (int k = 0; k < no; ++k) code |= (mt[oth[k]] + 8) << (4 + 4 * k);
                        m.hi_term.push_back(code);
                        m.hi_coef.push_back(c);
                    }
                    w.tend = static_cast<int32_t>(m.hi_term.size());
                    return w;
                };
                tt.hi_adj.push_back(emit_terms(false));
                if (p == 0 && pslot >= 0) tt.hi_feat.push_back(emit_terms(true));
            }
            continue;
        }
        // orders 3 and 4: one visit per factor position
        for (int p = 0; p < q; ++p) {
            std::vector<int> oth;
            for (int k = 0; k < q; ++k)
                if (k != p) oth.push_back(k);
            // other factors ordered by angular momentum: the product is
            // symmetric in them, and the unrolled kernel expects l_a <= l_b <= l_c
            std::stable_sort(oth.begin(), oth.end(), [&](int x, int y) { return ir.l[x] < ir.l[y]; });
            const int L = ir.l[p];
            VisitQ34 v{};
            v.order = q;
            v.itc = i;
            v.la = ir.l[oth[0]];
            v.sa = sp[oth[0]];
            v.lb = ir.l[oth[1]];
            v.sb = sp[oth[1]];
            v.lc = q == 4 ? ir.l[oth[2]] : 0;
            v.sc = q == 4 ? sp[oth[2]] : 0;
            v.pslot = pslot;
            const bool is_poly = pslot >= 0;
            int lmax_v = std::max(L, std::max(v.la, v.lb));
            if (q == 4) lmax_v = std::max(lmax_v, v.lc);
            // a visit runs unrolled only if the order-3/4 kernel was compiled
            // with its signature; any other goes to the general loop
            const bool fast = ((q == 3 && lmax_v <= kUnrollCap3) || (q == 4 && lmax_v <= kUnrollCap4)) &&
                              unrolled_signature(q, L, v.la, v.lb, v.lc);
            // adjoint visits accumulate G only; the energy of linear-only
            // families follows from G by Euler's theorem in the kernel
            v.flags = 2 | (fast ? 4 : 0);
            std::vector<int> mt(q);
            long cost = 0;
            auto emit = [&](bool dense_d, bool g_only_sum) {
                // block over (ia[, ib], k); k <-> m_p = -k.  Unrolled visits
                // keep only the valid terms, in loop order; the general loop
                // reads a dense block with zero padding.
                const int nA = 2 * v.la + 1, nB = q == 4 ? 2 * v.lb + 1 : 1;
                for (int ia = 0; ia < nA; ++ia)
                    for (int ib = 0; ib < nB; ++ib)
                        for (int k = 0; k <= L; ++k) {
                            const int mp = -k, ma = ia - v.la;
                            int mb, mc = 0;
                            bool ok;
                            if (q == 3) {
                                mb = -mp - ma;
                                ok = std::abs(mb) <= v.lb;
                            } else {
                                mb = ib - v.lb;
                                mc = -mp - ma - mb;
                                ok = std::abs(mc) <= v.lc;
                            }
                            double val = 0.0;
                            if (ok) {
                                std::fill(mt.begin(), mt.end(), 0);
                                mt[p] = mp;
                                mt[oth[0]] = ma;
                                mt[oth[1]] = mb;
                                if (q == 4) mt[oth[2]] = mc;
                                const double c = lookup(cm, mt);
                                if (g_only_sum) {
                                    val = c;
                                } else {
                                    std::vector<int> neg(mt);
                                    for (int& x : neg) x = -x;
                                    val = (k > 0 && dense_d) ? c + lookup(cm, neg) : c;
                                }
                                ++cost;
                            }
                            if (ok || !fast) push_coef(val);
                        }
            };
            v.coef = static_cast<int>(m.coef.size());
            emit(false, true);
            v.dcoef = -1;
            adj.push_back({sp[p], L, v, cost, static_cast<int>(m.coef.size()) - v.coef});
            if (p == 0 && is_poly) {
                VisitQ34 f = v;
                f.flags = 1 | (fast ? 4 : 0);
                f.dcoef = static_cast<int>(m.coef.size());
                f.coef = f.dcoef;
                emit(true, false);
                feat.push_back({sp[p], L, f, cost, static_cast<int>(m.coef.size()) - f.coef});
            }
        }
    }
    tt.q12_off.assign(m.maxl + 2, 0);
    for (int L = 0; L <= m.maxl; ++L) {
        tt.q12_off[L] = static_cast<int>(tt.visits_q12.size());
        tt.visits_q12.insert(tt.visits_q12.end(), by_l[L].begin(), by_l[L].end());
    }
    tt.q12_off[m.maxl + 1] = static_cast<int>(tt.visits_q12.size());
    // Factor positions of one family that carry the same (species, l) and
    // see the same other factors in the same order add the same products
    // into the same G row, term for term: merge them into one visit whose
    // coefficient block is the sum of theirs.
    {
        std::map<std::array<int, 11>, size_t> first;
        std::vector<Pending> merged;
        for (const Pending& q : adj) {
            const VisitQ34& v = q.v;
            const std::array<int, 11> key{q.S, q.L, v.itc, v.order, v.la, v.sa, v.lb, v.sb, v.lc, v.sc, v.flags};
            auto it = first.find(key);
            if (it == first.end()) {
                first.emplace(key, merged.size());
                merged.push_back(q);
                continue;
            }
            Pending& f = merged[it->second];
            for (int k = 0; k < q.clen; ++k) m.coef[f.v.coef + k] += m.coef[v.coef + k];
        }
        adj.swap(merged);
    }
    // every fused row is written by the order-3/4 kernel (the order-1/2
    // kernel skips it): rows no family adds to get an empty marker visit,
    // so their G is written as zero
    for (int S = 0; S < m.n_type; ++S)
        for (int L = 0; L <= std::max(m.q12_fuse, m.q12_fuse_hi); ++L) {
            if (!(L <= m.q12_fuse || (L > m.lmax3 && L <= m.q12_fuse_hi))) continue;
            bool any = false;
            for (const Pending& q : adj) any = any || (q.S == S && q.L == L);
            if (!any) {
                VisitQ34 z{};
                z.order = 0;
                z.flags = 2 | 4;
                z.pslot = -1;
                z.dcoef = -1;
                adj.push_back({S, L, z, 0, 0});
            }
        }
    tt.plan_adj = std::move(adj);
    tt.plan_feat = std::move(feat);
    m.types[t] = std::move(tt);
}

}  // namespace

// Radial-only style (des_type pair; mlip_model.h:45-101 without the gtinv
// lines, ModelParams mlip_model_params.cpp:27-82,172-221, PolynomialPair
// mlip_polynomial_pair.cpp:29-92): features d_{tc,n} = sum over neighbors of
// type combination tc of g_n(r); polynomial terms as the reference orders
// its coefficients.
bool load_pair_model(std::istream& in, Model& m, std::string& err, bool strict_count) {
    m.pair_model = true;
    if (m.model_type < 1 || m.model_type > 4 || m.maxp < 1 || m.maxp > 3) {
        err = "unsupported model_type/max_p";
        return false;
    }
    if (!(m.cutoff > 0.0)) {
        err = "cutoff must be positive";
        return false;
    }
    int nreg = 0;
    if (!first_token(in, nreg) || nreg <= 0) {
        err = "bad number of regression coefficients";
        return false;
    }
    if (!n_tokens(in, m.coeff_file, nreg) || !n_tokens(in, m.scale_file, nreg)) {
        err = "truncated coefficient or scale line";
        return false;
    }
    int npar = 0;
    if (!first_token(in, npar) || npar <= 0 || npar > kMaxFnPair) {
        err = "number of radial functions must be in [1,32]";
        return false;
    }
    m.beta.resize(npar);
    m.mu.resize(npar);
    for (int i = 0; i < npar; ++i) {
        std::vector<double> p;
        if (!n_tokens(in, p, 2)) {
            err = "truncated radial parameters";
            return false;
        }
        m.beta[i] = p[0];
        m.mu[i] = p[1];
    }
    if (!n_tokens(in, m.masses, m.n_type)) {
        err = "truncated mass line";
        return false;
    }
    if (m.n_type < 1 || m.n_type > 2) {
        err = "only one or two elements are supported (reference ModelParams)";
        return false;
    }
    m.n_fn = npar;
    m.tcp_other = m.n_type == 1 ? std::vector<std::array<int, 2>>{{0, -1}}
                                : std::vector<std::array<int, 2>>{{0, -1}, {1, 0}, {-1, 1}};
    m.n_tc = static_cast<int>(m.tcp_other.size());
    for (int t1 = 0; t1 < m.n_type; ++t1)
        for (int t2 = 0; t2 < m.n_type; ++t2)
            for (int tc = 0; tc < m.n_tc; ++tc)
                if (m.tcp_other[tc][t1] == t2) {
                    m.pair_tc[t1][t2] = tc;
                    break;
                }
    m.n_des = m.n_fn * m.n_tc;
    m.n_gtinv = 0;
    m.comb2.clear();
    m.comb3.clear();
    // a type combination is seen by centre type t1 if it lists a partner
    auto seen = [&](int tc, int t1) { return m.tcp_other[tc][t1] >= 0; };
    if (m.model_type == 1) {
        for (int f = 0; f < m.n_des; ++f) {
            if (m.maxp > 1) m.comb2.push_back({f, f});
            if (m.maxp > 2) m.comb3.push_back({f, f, f});
        }
        m.n_coeff = m.n_des * m.maxp;
    } else {
        // model_type 2 uses every feature; 3 and 4 none for this style
        std::vector<int> pidx;
        if (m.model_type == 2)
            for (int f = 0; f < m.n_des; ++f) pidx.push_back(f);
        auto tcof = [&](int f) { return f / m.n_fn; };
        if (m.maxp > 1)
            for (size_t i1 = 0; i1 < pidx.size(); ++i1)
                for (size_t i2 = 0; i2 <= i1; ++i2)
                    for (int t1 = 0; t1 < m.n_type; ++t1)
                        if (seen(tcof(pidx[i1]), t1) && seen(tcof(pidx[i2]), t1)) {
                            m.comb2.push_back({pidx[i2], pidx[i1]});
                            break;
                        }
        if (m.maxp > 2)
            for (size_t i1 = 0; i1 < pidx.size(); ++i1)
                for (size_t i2 = 0; i2 <= i1; ++i2)
                    for (size_t i3 = 0; i3 <= i2; ++i3)
                        for (int t1 = 0; t1 < m.n_type; ++t1)
                            if (seen(tcof(pidx[i1]), t1) && seen(tcof(pidx[i2]), t1) && seen(tcof(pidx[i3]), t1)) {
                                m.comb3.push_back({pidx[i3], pidx[i2], pidx[i1]});
                                break;
                            }
        m.n_coeff = m.n_des + static_cast<int>(m.comb2.size() + m.comb3.size());
    }
    if (!strict_count) return true;
    if (m.n_coeff != nreg) {
        char buf[200];
        std::snprintf(buf, sizeof buf, "potential lists %d coefficients, the model defines %d", nreg, m.n_coeff);
        err = buf;
        return false;
    }
    m.weight.resize(nreg);
    for (int i = 0; i < nreg; ++i) {
        if (m.scale_file[i] == 0.0) {
            err = "zero scale coefficient";
            return false;
        }
        m.weight[i] = m.coeff_file[i] / m.scale_file[i];
    }
    // per-feature derivative lists: every position of every product term
    std::vector<std::vector<PairHEntry>> lists(m.n_des);
    for (size_t k = 0; k < m.comb2.size(); ++k) {
        const double w = m.weight[m.n_des + k];
        const auto& c = m.comb2[k];
        lists[c[0]].push_back({c[1], -1, w, w / 2.0});
        lists[c[1]].push_back({c[0], -1, w, w / 2.0});
    }
    for (size_t k = 0; k < m.comb3.size(); ++k) {
        const double w = m.weight[m.n_des + m.comb2.size() + k];
        const auto& c = m.comb3[k];
        lists[c[0]].push_back({c[1], c[2], w, w / 3.0});
        lists[c[1]].push_back({c[0], c[2], w, w / 3.0});
        lists[c[2]].push_back({c[0], c[1], w, w / 3.0});
    }
    m.ph_off.assign(m.n_des + 1, 0);
    m.ph_ent.clear();
    for (int a = 0; a < m.n_des; ++a) {
        m.ph_ent.insert(m.ph_ent.end(), lists[a].begin(), lists[a].end());
        m.ph_off[a + 1] = static_cast<int32_t>(m.ph_ent.size());
    }
    // device layout: the feature rows reuse the polynomial-feature buffers
    m.n_p = m.n_des;
    m.maxl = 0;
    m.n_lmh = 1;
    m.n_lmh3 = 1;
    m.lmax3 = 0;
    m.has_order34 = false;
    m.has_feat34 = false;
    m.q12_fuse = -1;
    m.types.assign(m.n_type, TypeTables());
    return true;
}

bool load_model(const std::string& path, const GtinvTable& table, Model& m, std::string& err, bool strict_count) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open potential file " + path;
        return false;
    }
    std::string line;
    if (!next_line(in, line)) {
        err = "empty potential file";
        return false;
    }
    {
        std::istringstream ss(line);
        std::string tok;
        std::vector<std::string> toks;
        while (ss >> tok) toks.push_back(tok);
        // element names followed by exactly two trailing tokens (mlip_model.h:53-54)
        if (toks.size() < 3) {
            err = "element line must list elements followed by two comment tokens";
            return false;
        }
        m.elements.assign(toks.begin(), toks.end() - 2);
    }
    m.n_type = static_cast<int>(m.elements.size());
    std::string pair_type, des_type;
    if (!first_token(in, m.cutoff) || !first_token(in, pair_type) || !first_token(in, des_type) ||
        !first_token(in, m.model_type) || !first_token(in, m.maxp) || !first_token(in, m.maxl)) {
        err = "truncated potential header";
        return false;
    }
    if (pair_type != "gaussian") {
        err = "unknown pair_type " + pair_type + " (only gaussian)";
        return false;
    }
    if (des_type == "pair") return load_pair_model(in, m, err, strict_count);
    if (des_type != "gtinv") {
        err = "des_type must be gtinv or pair";
        return false;
    }
    if (m.model_type < 1 || m.model_type > 4 || m.maxp < 1 || m.maxp > 3) {
        err = "unsupported model_type/max_p";
        return false;
    }
    if (!(m.cutoff > 0.0) || m.maxl < 0 || m.maxl > kMaxL) {
        err = "cutoff must be positive and max_l in [0," + std::to_string(kMaxL) + "]";
        return false;
    }
    if (!first_token(in, m.gtinv_order) || m.gtinv_order < 1 || m.gtinv_order > 6) {
        err = "bad gtinv_order";
        return false;
    }
    if (!n_tokens(in, m.gtinv_maxl, m.gtinv_order - 1) || !n_tokens(in, m.gtinv_sym, m.gtinv_order - 1)) {
        err = "bad gtinv_maxl/gtinv_sym line";
        return false;
    }
    int nreg = 0;
    if (!first_token(in, nreg) || nreg <= 0) {
        err = "bad number of regression coefficients";
        return false;
    }
    if (!n_tokens(in, m.coeff_file, nreg) || !n_tokens(in, m.scale_file, nreg)) {
        err = "truncated coefficient or scale line";
        return false;
    }
    int npar = 0;
    if (!first_token(in, npar) || npar <= 0 || npar > kMaxFn) {
        err = "number of radial functions must be in [1," + std::to_string(kMaxFn) + "]";
        return false;
    }
    m.beta.resize(npar);
    m.mu.resize(npar);
    for (int i = 0; i < npar; ++i) {
        std::vector<double> p;
        if (!n_tokens(in, p, 2)) {
            err = "truncated radial parameters";
            return false;
        }
        m.beta[i] = p[0];
        m.mu[i] = p[1];
    }
    if (!n_tokens(in, m.masses, m.n_type)) {
        err = "truncated mass line";
        return false;
    }
    m.n_fn = npar;
    if (m.n_type < 1 || m.n_type > 2) {
        err = "only one or two elements are supported (reference ModelParams)";
        return false;
    }
    // type combinations (mlip_model_params.cpp:30-33)
    if (m.n_type == 1) {
        m.tcp_other = {{0, -1}};
    } else {
        m.tcp_other = {{0, -1}, {1, 0}, {-1, 1}};
    }
    m.n_tc = static_cast<int>(m.tcp_other.size());
    // screening (mlip_read_gtinv.cpp:35-76)
    m.irreps.clear();
    for (const Irrep& ir : table.irreps) {
        const int q = static_cast<int>(ir.l.size());
        if (q > 1) {
            if (q > m.gtinv_order || ir.l.back() > m.gtinv_maxl[q - 2]) continue;
            if (m.gtinv_sym[q - 2] && std::count(ir.l.begin(), ir.l.end(), ir.l[0]) != q) continue;
        }
        if (ir.l.back() > m.maxl) {
            err = "gtinv_maxl exceeds max_l";
            return false;
        }
        m.irreps.push_back(ir);
    }
    enumerate_itcs(m);
    enumerate_polynomial(m);
    if (!strict_count) return true;  // enumeration only (coefficient-count probe)
    if (m.n_coeff != nreg) {
        char buf[200];
        std::snprintf(buf, sizeof buf, "potential lists %d coefficients, the model defines %d", nreg, m.n_coeff);
        err = buf;
        return false;
    }
    m.weight.resize(nreg);
    for (int i = 0; i < nreg; ++i) {
        if (m.scale_file[i] == 0.0) {
            err = "zero scale coefficient";
            return false;
        }
        m.weight[i] = m.coeff_file[i] / m.scale_file[i];
    }
    for (const Irrep& ir : m.irreps) {
        const int q = static_cast<int>(ir.l.size());
        if (q > kMaxOrderGpu) {
            err = "invariants of order > 6 are not supported by the HIP path";
            return false;
        }
        if (q >= 5 && ir.l.back() > kMaxLHi) {
            err = "order-5/6 invariants with l > " + std::to_string(kMaxLHi) + " are not supported by the HIP path";
            return false;
        }
        if ((q == 3 || q == 4) && ir.l.back() > 15) {
            err = "order-3/4 invariants with l > 15 are not supported by the HIP path";
            return false;
        }
    }
    // polynomial features: per itc (all n or none)
    m.prank.assign(m.n_gtinv, -1);
    {
        std::vector<char> isp(m.n_gtinv, 0);
        for (int f : m.poly_features) isp[f % m.n_gtinv] = 1;
        // ranks grouped by the centre species that see the family: species 0
        // only, both, species 1 only, neither (coefficient order inside a
        // group), so each species' features are one contiguous rank range
        // and the quadratic form of a species' centres is a dense sub-block
        auto sees = [&](int i, int t) { return t < m.n_type && ((m.itcs[i].type_mask >> t) & 1u) != 0; };
        auto group = [&](int i) {
            const bool a = sees(i, 0), b = sees(i, 1);
            return a && !b ? 0 : (a && b ? 1 : (b ? 2 : 3));
        };
        int r = 0, cnt[4] = {0, 0, 0, 0};
        for (int gp = 0; gp < 4; ++gp)
            for (int i = 0; i < m.n_gtinv; ++i)
                if (isp[i] && group(i) == gp) {
                    m.prank[i] = r++;
                    ++cnt[gp];
                }
        m.n_pitc = r;
        m.prange[0][0] = 0;
        m.prange[0][1] = cnt[0] + cnt[1];
        m.prange[1][0] = cnt[0];
        m.prange[1][1] = cnt[0] + cnt[1] + cnt[2];
        m.n_p = m.n_fn * m.n_pitc;
        // every feature of a polynomial itc must be in the list for all n
        if (static_cast<int>(m.poly_features.size()) != m.n_p) {
            err = "internal: polynomial features are not closed over n";
            return false;
        }
    }
    auto pos = [&](int f) { return (f / m.n_gtinv) * m.n_pitc + m.prank[f % m.n_gtinv]; };
    m.lin_p.assign(m.n_p, 0.0);
    for (int f : m.poly_features) m.lin_p[pos(f)] = m.weight[f];
    m.wmat.assign(static_cast<size_t>(m.n_p) * m.n_p, 0.0);
    for (size_t k = 0; k < m.comb2.size(); ++k) {
        const double w = m.weight[m.n_des + k];
        const int a = pos(m.comb2[k][0]), b = pos(m.comb2[k][1]);
        if (a == b) {
            m.wmat[static_cast<size_t>(a) * m.n_p + a] += 2.0 * w;
        } else {
            m.wmat[static_cast<size_t>(a) * m.n_p + b] += w;
            m.wmat[static_cast<size_t>(b) * m.n_p + a] += w;
        }
    }
    m.c3_idx.clear();
    m.c3_w.clear();
    for (size_t k = 0; k < m.comb3.size(); ++k) {
        for (int q = 0; q < 3; ++q) m.c3_idx.push_back(pos(m.comb3[k][q]));
        m.c3_w.push_back(m.weight[m.n_des + m.comb2.size() + k]);
    }
    // device layout
    m.n_lmh = n_lmh_of(m.maxl);
    m.lmax3 = 0;
    m.has_order34 = false;
    for (const Irrep& ir : m.irreps)
        if (ir.l.size() >= 3) {
            m.has_order34 = true;
            m.lmax3 = std::max(m.lmax3, ir.l.back());
        }
    m.n_lmh3 = n_lmh_of(m.lmax3);
    // order-1/2 adjoint rows inside the order-3/4 kernel: those within its
    // LDS tile and unrolled signatures (l <= min(lmax3, 8)), and those above
    // the tile up to kMaxLQ12 (a_nlm read from HBM)
    m.q12_fuse = m.has_order34 ? std::min(m.lmax3, kUnrollCap3) : -1;
    m.q12_fuse_hi = m.has_order34 ? std::min(m.maxl, kMaxLQ12) : -1;
    m.lmaxh = -1;
    for (const Irrep& ir : m.irreps)
        if (ir.l.size() >= 5) m.lmaxh = std::max(m.lmaxh, ir.l.back());
    m.has_hi = m.lmaxh >= 0;
    m.hi_term.clear();
    m.hi_coef.clear();
    std::vector<std::map<std::vector<int>, double>> cmap(m.irreps.size());
    for (size_t i = 0; i < m.irreps.size(); ++i)
        for (size_t t = 0; t < m.irreps[i].m.size(); ++t) cmap[i][m.irreps[i].m[t]] = m.irreps[i].c[t];
    m.coef.clear();
    m.types.assign(m.n_type, TypeTables());
    for (int t = 0; t < m.n_type; ++t) build_type_tables(m, t, cmap);
    plan_q34(m, m.plan_region, m.plan_visit, m.plan_pair);
    m.has_feat34 = false;
    for (const auto& tt : m.types)
        if (!tt.visits_feat.empty()) m.has_feat34 = true;
    return true;
}

void plan_q34(Model& m, int region_cost, int visit_cost, int pair_lmax) {
    m.plan_region = region_cost;
    m.plan_visit = visit_cost;
    m.plan_pair = pair_lmax;
    m.q34_red = 0;
    m.q34_dyn_adj = m.q34_dyn_feat = false;
    for (TypeTables& tt : m.types) {
        tt.groups_adj.clear();
        tt.visits_adj.clear();
        tt.groups_feat.clear();
        tt.visits_feat.clear();
        int feat_red = 0;
        make_groups(m, tt.plan_adj, tt.groups_adj, tt.visits_adj, true, m.q34_red, m.q34_dyn_adj, region_cost,
                    visit_cost);
        make_groups(m, tt.plan_feat, tt.groups_feat, tt.visits_feat, false, feat_red, m.q34_dyn_feat, region_cost,
                    visit_cost);
    }
}

}  // namespace mlip
