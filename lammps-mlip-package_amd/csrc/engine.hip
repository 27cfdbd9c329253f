// Host engine and C ABI of the gtinv evaluation path: owns the model tables
// on the device, converts LAMMPS neighbor data into the device pair layout,
// runs the kernel pipeline on one HIP stream and returns E, F, virial.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <thread>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine_internal.h"
#include "neighbor.h"
#include "workload.h"

using namespace mlip;

namespace mlip {

bool hip_ok(mlip_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

}  // namespace mlip

namespace {

void set_err(char* buf, size_t len, const std::string& msg) {
    if (buf && len) std::snprintf(buf, len, "%s", msg.c_str());
}

// The order-3/4 workgroup plan of every centre type (groups and packed
// visit descriptors) into device tables, with their ranges in c->dm.
bool upload_q34_plan(mlip_ctx* c) {
    const Model& m = c->model;
    DevModel& dm = c->dm;
    std::vector<VisitQ34> vadj, vfeat;
    std::vector<GroupQ34> gadj, gfeat;
    for (int t = 0; t < m.n_type; ++t) {
        const TypeTables& tt = m.types[t];
        auto append = [](const std::vector<GroupQ34>& gs, const std::vector<VisitQ34>& vs,
                         std::vector<GroupQ34>& gout, std::vector<VisitQ34>& vout, int32_t& beg, int32_t& end) {
            const int vb = static_cast<int>(vout.size());
            beg = static_cast<int>(gout.size());
            for (GroupQ34 g : gs) {
                g.vbeg += vb;
                g.vend += vb;
                gout.push_back(g);
            }
            end = static_cast<int>(gout.size());
            vout.insert(vout.end(), vs.begin(), vs.end());
        };
        append(tt.groups_adj, tt.visits_adj, gadj, vadj, dm.gadj_beg[t], dm.gadj_end[t]);
        append(tt.groups_feat, tt.visits_feat, gfeat, vfeat, dm.gfeat_beg[t], dm.gfeat_end[t]);
    }
    // device form of the order-3/4 visits
    // adjoint visits: H sources h = pslot >= 0 ? pslot : -1 - itc, the row
    // of a paired group in flags bit0 (row 1) and bit3 (both rows)
    auto hsrc = [](int pslot, int itc) { return pslot >= 0 ? pslot : -1 - itc; };
    auto pack = [&](const std::vector<VisitQ34>& vs, std::vector<VisitPack>& out, bool adj) {
        for (const VisitQ34& v : vs) {
            int flags = v.flags;
            if (adj) {
                if (flags & (1 | 8)) return false;
                flags |= v.row == 1 ? 1 : v.row == 2 ? 8 : 0;
            }
            const int f[8] = {v.order, v.la, v.lb, v.lc, v.sa, v.sb, v.sc, flags};
            uint32_t sig = 0;
            for (int k = 0; k < 8; ++k) {
                if (f[k] < 0 || f[k] > 15) return false;
                sig |= static_cast<uint32_t>(f[k]) << (4 * k);
            }
            if (adj)
                out.push_back({static_cast<int32_t>(sig), v.coef, hsrc(v.pslot, v.itc),
                               v.row == 2 ? hsrc(v.pslot1, v.itc1) : 0});
            else
                out.push_back({static_cast<int32_t>(sig), v.coef, v.pslot, v.itc});
        }
        return true;
    };
    std::vector<VisitPack> vadjp, vfeatp;
    if (!pack(vadj, vadjp, true) || !pack(vfeat, vfeatp, false)) {
        c->err = "order-3/4 visit outside the packed range";
        return false;
    }
    if (!upload(c, c->d_vadj, vadjp.data(), vadjp.size(), "adjoint visits") ||
        !upload(c, c->d_vfeat, vfeatp.data(), vfeatp.size(), "feature visits") ||
        !upload(c, c->d_gadj, gadj.data(), gadj.size(), "adjoint groups") ||
        !upload(c, c->d_gfeat, gfeat.data(), gfeat.size(), "feature groups"))
        return false;
    dm.vadj = c->d_vadj.p;
    dm.gadj = c->d_gadj.p;
    dm.vfeat = c->d_vfeat.p;
    dm.gfeat = c->d_gfeat.p;
    dm.q34_red = m.q34_red;
    dm.q34_dyn_adj = m.q34_dyn_adj ? 1 : 0;
    dm.q34_dyn_feat = m.q34_dyn_feat ? 1 : 0;
    return true;
}

bool upload_model(mlip_ctx* c) {
    const Model& m = c->model;
    DevModel& dm = c->dm;
    std::memset(&dm, 0, sizeof dm);
    dm.n_type = m.n_type;
    dm.n_fn = m.n_fn;
    dm.n_gtinv = m.n_gtinv;
    dm.n_lmh = m.n_lmh;
    dm.n_lmh3 = m.n_lmh3;
    dm.maxl = m.maxl;
    dm.lmax3 = m.lmax3;
    dm.n_pitc = m.n_pitc;
    dm.n_p = m.n_p;
    for (int t = 0; t < 2; ++t)
        for (int k = 0; k < 2; ++k) dm.prange[t][k] = m.prange[t][k];
    dm.has_q34 = m.has_order34 ? 1 : 0;
    dm.q12_fuse = m.q12_fuse;
    dm.q12_fuse_hi = m.q12_fuse_hi;
    for (int L = 0; L <= m.maxl; ++L)
        if (!q12_fused(dm, L)) dm.q12_adj_levels = 1;
    dm.q34_red = m.q34_red;
    dm.pair_model = m.pair_model ? 1 : 0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) dm.pair_tc[a][b] = m.pair_tc[a][b];
    dm.q34_dyn_adj = m.q34_dyn_adj ? 1 : 0;
    dm.q34_dyn_feat = m.q34_dyn_feat ? 1 : 0;
    dm.cutoff = m.cutoff;
    for (int n = 0; n < m.n_fn; ++n) {
        dm.beta[n] = m.beta[n];
        dm.mu[n] = m.mu[n];
    }
    std::vector<VisitQ12> q12;
    std::vector<int32_t> q12off;
    std::vector<FeatQ12> fq12;
    for (int t = 0; t < m.n_type; ++t) {
        const TypeTables& tt = m.types[t];
        const int base = static_cast<int>(q12.size());
        for (int v : tt.q12_off) q12off.push_back(base + v);
        q12.insert(q12.end(), tt.visits_q12.begin(), tt.visits_q12.end());
        dm.fq12_off[t] = static_cast<int32_t>(fq12.size());
        // (the radial-only style has no order-1/2 tables: q12_off is empty)
        const int nl = static_cast<int>(tt.q12_off.size()) - 1;
        for (int L = 0; L < nl && L <= m.maxl; ++L)
            for (int k = tt.q12_off[L]; k < tt.q12_off[L + 1]; ++k) {
                const VisitQ12& v = tt.visits_q12[k];
                if (v.pslot >= 0) fq12.push_back({v.order, L, v.s0, v.s1, v.pslot - m.prange[t][0], v.coef});
            }
    }
    for (int t = m.n_type; t < 3; ++t) dm.fq12_off[t] = static_cast<int32_t>(fq12.size());
    if (!upload_q34_plan(c)) return false;
    // Polynomial features per species, compact: species t keeps the ranks
    // [lo, hi) = prange[t] it sees (the ranks are grouped so that this is
    // one range), column ci(n, r) = n (hi - lo) + r - lo, in a slot row of
    // n_pr doubles; the model's tables use the feature n * n_pitc + r.
    int n_pr = 0;
    for (int t = 0; t < m.n_type && t < 2; ++t)
        n_pr = std::max(n_pr, m.n_fn * (m.prange[t][1] - m.prange[t][0]));
    n_pr = (n_pr + 15) / 16 * 16;  // rows start on 128-byte lines
    if (n_pr > m.n_p) n_pr = m.n_p;  // the row buffers hold n_p per slot
    dm.n_pr = n_pr;
    // feature ranks of the range no visit of a species writes (zeroed by
    // k_q12<true>), as offsets r - lo
    std::vector<int32_t> zlist;
    for (int t = 0; t < m.n_type && t < 2; ++t) {
        std::vector<char> written(static_cast<size_t>(m.n_pitc), 0);
        for (const VisitQ12& v : m.types[t].visits_q12)
            if (v.pslot >= 0) written[v.pslot] = 1;
        for (const VisitQ34& v : m.types[t].visits_feat)
            if (v.pslot >= 0) written[v.pslot] = 1;
        for (const VisitHi& v : m.types[t].hi_feat)
            if (v.pslot >= 0) written[v.pslot] = 1;
        dm.zoff[t] = static_cast<int32_t>(zlist.size());
        for (int r = 0; r < m.n_pitc; ++r) {
            const bool in = r >= m.prange[t][0] && r < m.prange[t][1];
            if (written[r] && !in) {
                c->err = "internal: polynomial feature outside its species' rank range";
                return false;
            }
            if (in && !written[r]) zlist.push_back(r - m.prange[t][0]);
        }
    }
    for (int t = m.n_type; t < 2; ++t) dm.zoff[t] = static_cast<int32_t>(zlist.size());
    dm.zoff[2] = static_cast<int32_t>(zlist.size());
    // per species: quadratic form on its columns (dense, for one GEMM per
    // species), linear weights, cubic terms transposed by target column
    // (terms with a factor the species does not see vanish and are dropped)
    std::vector<double> wc, linc(static_cast<size_t>(2) * n_pr, 0.0), c3w;
    std::vector<int32_t> pdeg(static_cast<size_t>(2) * n_pr, 1);
    std::vector<int> order_of_rank(static_cast<size_t>(std::max(m.n_pitc, 1)), 1);
    for (int i = 0; i < m.n_gtinv && !m.pair_model; ++i)
        if (m.prank[i] >= 0) order_of_rank[m.prank[i]] = m.itcs[i].order;
    std::vector<int32_t> c3off(static_cast<size_t>(2) * (n_pr + 1), 0), c3bc;
    for (int t = 0; t < 2; ++t) {
        const bool on = t < m.n_type && m.n_p > 0;
        const int lo = m.prange[t][0], wr = on ? m.prange[t][1] - lo : 0, wf = m.n_fn * wr;
        c->wofs[t] = wc.size();
        c->wdim[t] = wf;
        std::vector<int> gof(static_cast<size_t>(wf));  // compact column -> model feature
        std::vector<int> cof(static_cast<size_t>(m.n_p), -1);  // model feature -> compact column
        for (int ci = 0; ci < wf; ++ci) {
            gof[ci] = (ci / wr) * m.n_pitc + lo + ci % wr;
            cof[gof[ci]] = ci;
        }
        for (int ci = 0; ci < wf; ++ci)
            for (int cj = 0; cj < wf; ++cj)
                wc.push_back(m.wmat[static_cast<size_t>(gof[ci]) * m.n_p + gof[cj]]);
        for (int ci = 0; ci < wf; ++ci) {
            linc[static_cast<size_t>(t) * n_pr + ci] = m.lin_p[gof[ci]];
            pdeg[static_cast<size_t>(t) * n_pr + ci] = order_of_rank[gof[ci] % m.n_pitc];
        }
        std::vector<std::vector<std::pair<std::array<int, 2>, double>>> c3(static_cast<size_t>(wf));
        for (size_t k = 0; on && k < m.c3_w.size(); ++k) {
            const int a = cof[m.c3_idx[3 * k]], b = cof[m.c3_idx[3 * k + 1]], cc = cof[m.c3_idx[3 * k + 2]];
            if (a < 0 || b < 0 || cc < 0) continue;
            const double w = m.c3_w[k];
            c3[a].push_back({{b, cc}, w});
            c3[b].push_back({{a, cc}, w});
            c3[cc].push_back({{a, b}, w});
        }
        int32_t* off = c3off.data() + static_cast<size_t>(t) * (n_pr + 1);
        for (int ci = 0; ci <= n_pr; ++ci) {
            off[ci] = static_cast<int32_t>(c3w.size());
            if (ci < wf)
                for (const auto& e : c3[ci]) {
                    c3bc.push_back(e.first[0]);
                    c3bc.push_back(e.first[1]);
                    c3w.push_back(e.second);
                }
        }
    }
    // linear weights of the per-atom energy split: w / order of the family
    std::vector<double> wsplit(m.weight.begin(), m.weight.end());
    if (!m.pair_model)
        for (int f = 0; f < m.n_des && f < static_cast<int>(wsplit.size()); ++f)
            wsplit[f] /= m.itcs[f % m.n_gtinv].order;
    bool ok = upload(c, c->d_coef, m.coef.data(), m.coef.size(), "coefficients") &&
              upload(c, c->d_weight_e, wsplit.data(), wsplit.size(), "split weights") &&
              upload(c, c->d_pdeg, pdeg.data(), pdeg.size(), "feature degrees") &&
              upload(c, c->d_weight, m.weight.data(), m.weight.size(), "weights") &&
              upload(c, c->d_wmat, wc.data(), wc.size(), "quadratic form") &&
              upload(c, c->d_linp, linc.data(), linc.size(), "linear poly weights") &&
              upload(c, c->d_c3w, c3w.data(), c3w.size(), "cubic weights") &&
              upload(c, c->d_c3bc, c3bc.data(), c3bc.size(), "cubic index") &&
              upload(c, c->d_c3off, c3off.data(), c3off.size(), "cubic offsets") &&
              upload(c, c->d_q12, q12.data(), q12.size(), "q12 visits") &&
              upload(c, c->d_q12off, q12off.data(), q12off.size(), "q12 offsets") &&
              upload(c, c->d_fq12, fq12.data(), fq12.size(), "order-1/2 feature list") &&
              upload(c, c->d_zlist, zlist.data(), zlist.size(), "feature zero list");
    if (!ok) return false;
    dm.poly_fused = (!m.pair_model && m.n_p > 0 && c3w.empty()) ? 1 : 0;
    dm.zlist = c->d_zlist.p;
    dm.fq12 = c->d_fq12.p;
    if (m.pair_model) {
        if (!upload(c, c->d_pairw, m.weight.data(), static_cast<size_t>(m.n_des), "pair weights") ||
            !upload(c, c->d_phoff, m.ph_off.data(), m.ph_off.size(), "pair derivative offsets") ||
            !upload(c, c->d_phent, m.ph_ent.data(), m.ph_ent.size(), "pair derivative lists"))
            return false;
        dm.pair_w = c->d_pairw.p;
        dm.ph_off = c->d_phoff.p;
        dm.ph_ent = c->d_phent.p;
    }
    if (m.has_hi) {
        std::vector<VisitHi> hv;
        for (int t = 0; t < m.n_type && t < 2; ++t) {
            dm.hadj_beg[t] = static_cast<int32_t>(hv.size());
            hv.insert(hv.end(), m.types[t].hi_adj.begin(), m.types[t].hi_adj.end());
            dm.hadj_end[t] = static_cast<int32_t>(hv.size());
            dm.hfeat_beg[t] = static_cast<int32_t>(hv.size());
            hv.insert(hv.end(), m.types[t].hi_feat.begin(), m.types[t].hi_feat.end());
            dm.hfeat_end[t] = static_cast<int32_t>(hv.size());
        }
        if (!upload(c, c->d_hvis, hv.data(), hv.size(), "order-5/6 visits") ||
            !upload(c, c->d_hterm, m.hi_term.data(), m.hi_term.size(), "order-5/6 terms") ||
            !upload(c, c->d_hcoef, m.hi_coef.data(), m.hi_coef.size(), "order-5/6 coefficients"))
            return false;
        dm.hvis = c->d_hvis.p;
        dm.hterm = c->d_hterm.p;
        dm.hcoef = c->d_hcoef.p;
        dm.lmaxh = m.lmaxh;
        dm.n_lmhh = n_lmh_of(m.lmaxh);
    } else {
        dm.lmaxh = -1;
        dm.n_lmhh = 0;
    }
    {
        YTab<double> t64;
        YTab<float> t32;
        build_ylm_tables(t64, t32);
        if (!upload(c, c->d_yt64, &t64, 1, "Y_lm recursion table") ||
            !upload(c, c->d_yt32, &t32, 1, "Y_lm recursion table (FP32)"))
            return false;
    }
    dm.yt64 = c->d_yt64.p;
    dm.yt32 = c->d_yt32.p;
    dm.coef = c->d_coef.p;
    dm.weight = c->d_weight.p;
    dm.q12 = c->d_q12.p;
    dm.q12_off = c->d_q12off.p;
    dm.lin_p = c->d_linp.p;
    dm.c3_off = c->d_c3off.p;
    dm.c3_bc = c->d_c3bc.p;
    dm.c3_w = c->d_c3w.p;
    dm.weight_e = c->d_weight_e.p;
    dm.pdeg = c->d_pdeg.p;
    dm.wmat = c->d_wmat.p;
    for (int t = 0; t < 2; ++t) {
        dm.wofs[t] = static_cast<int32_t>(c->wofs[t]);
        dm.wdim[t] = c->wdim[t];
    }
    return hip_ok(c, hipStreamSynchronize(c->stream), "model upload");
}

}  // namespace

namespace mlip {

// Slot assignment: centres sorted by element (stable), each element padded
// to whole blocks of kBlock lanes.
void assign_slots(const mlip_ctx* c, int nown, const std::vector<int>& elem_of_owner,
                  std::vector<int32_t>& atom_of_slot, std::vector<int32_t>& slot_of_atom,
                  std::vector<int32_t>& block_type) {
    atom_of_slot.clear();
    block_type.clear();
    slot_of_atom.assign(nown, -1);
    for (int t = 0; t < c->model.n_type; ++t) {
        std::vector<int> mem;
        for (int i = 0; i < nown; ++i)
            if (elem_of_owner[i] == t) mem.push_back(i);
        const int nb = static_cast<int>((mem.size() + kBlock - 1) / kBlock);
        for (int b = 0; b < nb; ++b) {
            block_type.push_back(t);
            for (int l = 0; l < kBlock; ++l) {
                const size_t k = static_cast<size_t>(b) * kBlock + l;
                const int a = k < mem.size() ? mem[k] : -1;
                if (a >= 0) slot_of_atom[a] = static_cast<int>(atom_of_slot.size());
                atom_of_slot.push_back(a);
            }
        }
    }
}

}  // namespace mlip

namespace {

// Slot layout, work buffers and the DevSystem view of a structure whose
// positions, elements, pair lists and gather lists are already on the device.
bool setup_layout(mlip_ctx* c, int nall, int nown, int64_t npairs, const std::vector<int32_t>& elem_own, bool has_pcen,
                  bool has_shift) {
    const Model& m = c->model;
    drop_graph(c);
    std::vector<int32_t> atom_of_slot, slot_of_atom, block_type;
    assign_slots(c, nown, elem_own, atom_of_slot, slot_of_atom, block_type);
    const WorkSizes z = work_sizes(m, nall, nown, npairs, static_cast<int>(block_type.size()));
    if (!c->atom_of_slot.ensure(z.nslot) || !c->slot_of_atom.ensure(nown) || !c->block_type.ensure(z.nblocks) ||
        !c->A.ensure(z.asz) || !c->G.ensure(z.asz) || !c->D.ensure(z.psz) || !c->Hq.ensure(z.psz) ||
        !c->H.ensure(z.psz) || !c->Ep.ensure(z.nslot) || !c->Eq12.ensure(z.nslot * m.n_fn) ||
        !c->Eq34.ensure(z.nslot * m.n_fn * kWavesPerWg) || !c->gpair.ensure(static_cast<size_t>(npairs) * 3) ||
        !c->vown.ensure(static_cast<size_t>(nown) * 6) || !c->f.ensure(static_cast<size_t>(nall) * 3) ||
        !c->e_slot.ensure(z.nslot) || !c->red.ensure(7 * 128) || !c->out.ensure(8)) {
        c->err = "device allocation failed for structure buffers";
        return false;
    }
    SysBufs b{};
    b.x = c->x.p;
    b.pshift = has_shift ? c->pshift.p : nullptr;
    b.elem = c->elem.p;
    b.poff = c->poff.p;
    b.pnbr = c->pnbr.p;
    b.pcen = has_pcen ? c->pcen.p : nullptr;
    b.goff = c->goff.p;
    b.gent = c->gent.p;
    b.atom_of_slot = c->atom_of_slot.p;
    b.slot_of_atom = c->slot_of_atom.p;
    b.block_type = c->block_type.p;
    b.A = c->A.p;
    b.G = c->G.p;
    b.D = c->D.p;
    b.Hq = c->Hq.p;
    b.H = c->H.p;
    b.Ep = c->Ep.p;
    b.Eq12 = c->Eq12.p;
    b.Eq34 = c->Eq34.p;
    b.gpair = c->gpair.p;
    b.vown = c->vown.p;
    b.f = c->f.p;
    b.e_slot = c->e_slot.p;
    b.red = c->red.p;
    b.out = c->out.p;
    return bind_system(c, nall, nown, npairs, elem_own, atom_of_slot, slot_of_atom, block_type, b);
}

// Upload a structure with host-built lists and lay it out.
bool setup_system(mlip_ctx* c, int nall, int nown, int64_t npairs, const std::vector<double>& x,
                  const std::vector<int32_t>& elem, const std::vector<int32_t>& elem_own,
                  const std::vector<int32_t>& poff,
                  const std::vector<int32_t>& pnbr, const std::vector<int32_t>* pcen,
                  const std::vector<double>* pshift, const std::vector<int32_t>& goff,
                  const std::vector<int32_t>& gent) {
    bool ok = upload(c, c->x, x.data(), x.size(), "positions") &&
              upload(c, c->elem, elem.data(), elem.size(), "elements") &&
              upload(c, c->poff, poff.data(), poff.size(), "pair offsets") &&
              upload(c, c->pnbr, pnbr.data(), pnbr.size(), "pair neighbors") &&
              upload(c, c->goff, goff.data(), goff.size(), "gather offsets") &&
              upload(c, c->gent, gent.data(), gent.size(), "gather entries");
    if (!ok) return false;
    if (pcen && !upload(c, c->pcen, pcen->data(), pcen->size(), "pair centres")) return false;
    if (pshift && !upload(c, c->pshift, pshift->data(), pshift->size(), "pair shifts")) return false;
    return setup_layout(c, nall, nown, npairs, elem_own, pcen != nullptr, pshift != nullptr);
}

// Structures from periodic cells with their lists built on the device
// (positions concatenated, natoms[s] each; elem per atom).  Returns an
// MLIP_* code: invalid cells or positions are argument errors.
int setup_cells_device(mlip_ctx* c, int nstruct, const int32_t* natoms, const double* cells, const double* x,
                       const std::vector<int32_t>& elem) {
    const size_t n = elem.size();
    for (size_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(x[i])) {
            c->err = "non-finite position";
            return MLIP_ERR_ARG;
        }
    for (int k = 0; k < nstruct; ++k) {
        CellBins cb;
        std::string e;
        if (natoms[k] > 0 && !cell_bins(cells + 9 * static_cast<size_t>(k), c->model.cutoff, cb, e)) {
            c->err = "structure " + std::to_string(k) + ": " + e;
            return MLIP_ERR_ARG;
        }
    }
    int64_t npairs = 0;
    if (!upload(c, c->x, x, 3 * n, "positions") || !upload(c, c->elem, elem.data(), n, "elements") ||
        !device_lists(c, nstruct, natoms, cells, c->model.cutoff, npairs))
        return MLIP_ERR_DEVICE;
    return setup_layout(c, static_cast<int>(n), static_cast<int>(n), npairs, elem, false, true) ? MLIP_OK
                                                                                                  : MLIP_ERR_DEVICE;
}

// Per-structure atom ranges and result slots of a batch; synchronises.
int finish_batch(mlip_ctx* c, int nstruct, const std::vector<int32_t>& soff, int n) {
    if (!upload(c, c->soff, soff.data(), soff.size(), "structure offsets") || !c->sout.ensure(7 * nstruct))
        return MLIP_ERR_DEVICE;
    if (!hip_ok(c, hipStreamSynchronize(c->stream), "batch upload")) return MLIP_ERR_DEVICE;
    c->nstruct = nstruct;
    c->sbeg = c->soff.p;
    c->send = c->soff.p + 1;
    c->sres = c->sout.p;
    c->resident = true;
    c->resident_natoms = n;
    return MLIP_OK;
}

}  // namespace

namespace mlip {

// Force gather lists: for every position k, the pairs whose centre (-) or
// neighbor (+) is k, in increasing pair order.
void build_gather(int nall, const std::vector<int32_t>& poff, const std::vector<int32_t>& pnbr,
                  const std::vector<int32_t>* pcen, std::vector<int32_t>& goff, std::vector<int32_t>& gent) {
    const int nown = static_cast<int>(poff.size()) - 1;
    const size_t np = pnbr.size();
    std::vector<int32_t> cnt(nall + 1, 0);
    std::vector<int32_t> cen(np);
    for (int o = 0; o < nown; ++o)
        for (int p = poff[o]; p < poff[o + 1]; ++p) cen[p] = pcen ? (*pcen)[p] : o;
    for (size_t p = 0; p < np; ++p) {
        ++cnt[cen[p] + 1];
        ++cnt[pnbr[p] + 1];
    }
    goff.assign(nall + 1, 0);
    for (int k = 0; k < nall; ++k) goff[k + 1] = goff[k] + cnt[k + 1];
    gent.assign(goff[nall], 0);
    std::vector<int32_t> fill(goff.begin(), goff.end() - 1);
    for (size_t p = 0; p < np; ++p) {
        gent[fill[cen[p]]++] = -static_cast<int32_t>(p + 1);
        gent[fill[pnbr[p]]++] = static_cast<int32_t>(p + 1);
    }
}

WorkSizes work_sizes(const Model& m, int nall, int nown, int64_t npairs, int nblocks) {
    WorkSizes z;
    z.nblocks = static_cast<size_t>(nblocks);
    z.nslot = z.nblocks * kBlock;
    z.asz = z.nslot * m.n_fn * m.n_type * m.n_lmh * 2;
    z.psz = z.nslot * static_cast<size_t>(m.n_p);
    (void)nall;
    (void)nown;
    (void)npairs;
    return z;
}

// Slot arrays into b (copied on the stream from host vectors that must live
// until the stream has passed them: the caller synchronises before it
// returns), a_nlm cleared, and the DevSystem view of the structure.
bool bind_system(mlip_ctx* c, int nall, int nown, int64_t npairs, const std::vector<int32_t>& elem_own,
                 const std::vector<int32_t>& atom_of_slot, const std::vector<int32_t>& slot_of_atom,
                 const std::vector<int32_t>& block_type, const SysBufs& b) {
    const Model& m = c->model;
    c->nstruct = 0;
    c->sres = nullptr;
    c->total_reduce = true;
    const int nblocks = static_cast<int>(block_type.size());
    c->nblk_type[0] = c->nblk_type[1] = 0;
    for (int32_t t : block_type) ++c->nblk_type[t];
    c->owners_per_type.assign(m.n_type, 0);
    for (int e : elem_own)
        if (e >= 0 && e < m.n_type) ++c->owners_per_type[e];
    c->npairs = npairs;
    const WorkSizes z = work_sizes(m, nall, nown, npairs, nblocks);
    hipStream_t st = c->stream;
    bool ok = hip_ok(c, hipMemcpyAsync(b.atom_of_slot, atom_of_slot.data(), z.nslot * sizeof(int32_t),
                                       hipMemcpyHostToDevice, st), "slots") &&
              hip_ok(c, hipMemcpyAsync(b.slot_of_atom, slot_of_atom.data(), nown * sizeof(int32_t),
                                       hipMemcpyHostToDevice, st), "slots") &&
              hip_ok(c, hipMemcpyAsync(b.block_type, block_type.data(), z.nblocks * sizeof(int32_t),
                                       hipMemcpyHostToDevice, st), "block types") &&
              hip_ok(c, launch_fill(b.A, z.asz, st), "clear a_nlm") &&
              hip_ok(c, launch_fill(b.vown, static_cast<size_t>(nown) * 6, st), "clear virial");
    if (!ok) return false;
    DevSystem& s = c->ds;
    s = DevSystem{};
    s.nall = nall;
    s.nown = nown;
    s.nblocks = nblocks;
    s.npairs = npairs;
    s.x = b.x;
    s.elem = b.elem;
    s.poff = b.poff;
    s.pend = b.poff + 1;
    s.pnbr = b.pnbr;
    s.pcen = b.pcen;
    s.pshift = b.pshift;
    s.atom_of_slot = b.atom_of_slot;
    s.slot_of_atom = b.slot_of_atom;
    s.block_type = b.block_type;
    s.A = b.A;
    s.G = b.G;
    s.D = b.D;
    s.Hq = b.Hq;
    s.H = b.H;
    s.Ep = b.Ep;
    s.Eq12 = b.Eq12;
    s.Eq34 = b.Eq34;
    s.gpair = b.gpair;
    s.vown = b.vown;
    s.goff = b.goff;
    s.gend = b.goff + 1;
    s.gent = b.gent;
    s.f = b.f;
    s.e_slot = b.e_slot;
    s.red = b.red;
    s.out = b.out;
    return true;
}

}  // namespace mlip

namespace {

void prof_mark(mlip_ctx* c, const char* name, bool start) {
    if (!c->prof) return;
    if (start) {
        if (c->ev_used == c->events.size()) {
            ProfEvent e{name, nullptr, nullptr};
            (void)hipEventCreate(&e.a);
            (void)hipEventCreate(&e.b);
            c->events.push_back(e);
        }
        c->events[c->ev_used].name = name;
        (void)hipEventRecord(c->events[c->ev_used].a, c->stream);
    } else {
        (void)hipEventRecord(c->events[c->ev_used].b, c->stream);
        ++c->ev_used;
    }
}

struct Scope {
    mlip_ctx* c;
    Scope(mlip_ctx* cc, const char* n) : c(cc) { prof_mark(c, n, true); }
    ~Scope() { prof_mark(c, nullptr, false); }
};

}  // namespace

namespace mlip {

bool run_pipeline(mlip_ctx* c) {
    const DevModel& m = c->dm;
    const DevSystem& s = c->ds;
    const Model& mm = c->model;
    hipStream_t st = c->stream;
    // the first failing launch names the call's error; later launches are
    // not enqueued after one failed
    bool ok = true;
    auto run = [&](const char* name, auto&& launch) {
        if (!ok) return;
        Scope sc(c, name);
        const hipError_t e = launch();
        if (e != hipSuccess) {
            ok = false;
            c->err = std::string("kernel launch ") + name + ": " + hipGetErrorString(e);
        }
    };
    if (mm.pair_model) {
        // radial-only style: features, polynomial and H, pair forces
        run("pair_features", [&] { return launch_pair_features(m, s, st); });
        run("pair_poly", [&] { return launch_pair_poly(m, s, st); });
        run("force", [&] { return launch_pair_force(m, s, st); });
        run("gather", [&] { return launch_gather(m, s, st); });
        run("energy_virial", [&] { return launch_energy_virial(m, s, c->total_reduce, st); });
        if (c->nstruct > 0)
            run("struct_reduce", [&] { return launch_struct_reduce(s, c->sbeg, c->send, c->nstruct, c->sres, st); });
        return ok;
    }
    // the register a_nlm kernel forms the order-1/2 features itself where it can
    const bool ff = m.n_p > 0 && anlm_forms_features(m, c->precision, c->anlm_kind);
    run("anlm", [&] { return launch_anlm(m, s, c->precision, st, c->anlm_kind, ff); });
    if (m.n_p > 0) {
        if (!ff) run("q12_feat", [&] { return launch_q12(m, s, true, st); });
        if (mm.has_feat34) run("q34_feat", [&] { return launch_q34_feat(m, s, st); });
        if (mm.has_hi) run("qhigh_feat", [&] { return launch_qhigh(m, s, true, st); });
        if (!ok) return false;
        // Hq = D W per species: slot blocks are grouped by species and a
        // species' row holds only the features it sees, so each species is
        // one dense product on its rows and its own quadratic form.  Without
        // cubic terms the product's epilogue also forms H and the energy
        // partials, and "poly" only sums those; otherwise k_poly does both
        run("poly_gemm", [&] { return launch_poly_gemm(m, s, c->nblk_type, st); });
        run("poly", [&] { return launch_poly(m, s, st); });
    }
    run("q12_adj", [&] { return launch_q12(m, s, false, st); });
    if (mm.has_order34) run("q34_adj", [&] { return launch_q34_adj(m, s, st); });
    if (mm.has_hi) run("qhigh_adj", [&] { return launch_qhigh(m, s, false, st); });
    run("force", [&] { return launch_force(m, s, c->precision, c->force_threads, st); });
    run("gather", [&] { return launch_gather(m, s, st); });
    run("energy_virial", [&] { return launch_energy_virial(m, s, c->total_reduce, st); });
    if (c->nstruct > 0)
        run("struct_reduce", [&] { return launch_struct_reduce(s, c->sbeg, c->send, c->nstruct, c->sres, st); });
    return ok;
}

void drop_graph(mlip_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->graph_warm = false;
}

// One evaluation: replays the captured graph when graph mode is on.  After a
// structure or setting change the first call runs eagerly (kernel attributes
// such as the dynamic LDS limits are set outside any capture) and the second
// captures.  Profiling events are not recorded inside a capture.
bool run_eval(mlip_ctx* c) {
    if (!c->use_graph || c->prof) return run_pipeline(c);
    if (!c->gexec && !c->graph_warm) {
        c->graph_warm = true;
        return run_pipeline(c);
    }
    if (!c->gexec) {
        hipGraph_t g = nullptr;
        if (!hip_ok(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture"))
            return false;
        const bool ok = run_pipeline(c);
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (!ok || !hip_ok(c, e, "end capture")) {
            if (g) (void)hipGraphDestroy(g);
            return false;
        }
        const hipError_t ie = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (!hip_ok(c, ie, "graph instantiate")) {
            c->gexec = nullptr;
            return false;
        }
    }
    return hip_ok(c, hipGraphLaunch(c->gexec, c->stream), "graph launch");
}

// Per-atom energy and virial in the reference's pair split, after an
// evaluation: the split adjoint (every polynomial term of dE/da weighted by
// 1 / its degree, linear families by 1 / their order) overwrites G, then
// per-pair shares (epv [npairs][7]) and per-position half sums (evat
// [nall][7]).
bool run_atom_split(mlip_ctx* c, double* epv, double* evat) {
    const Model& mm = c->model;
    DevModel me = c->dm;
    DevSystem se = c->ds;
    hipStream_t st = c->stream;
    bool ok = true;
    auto run = [&](const char* name, hipError_t e) {
        if (ok && e != hipSuccess) {
            ok = false;
            c->err = std::string("kernel launch ") + name + ": " + hipGetErrorString(e);
        }
        return ok;
    };
    if (mm.pair_model) {
        run("pair_model_energy", launch_pair_model_energy(me, se, epv, st));
    } else {
        me.weight = c->d_weight_e.p;
        se.H = se.Hq;
        if (me.n_p > 0) run("poly_split", launch_poly_split(me, se, st));
        if (ok) run("q12_split", launch_q12(me, se, false, st));
        if (ok && mm.has_order34) run("q34_split", launch_q34_adj(me, se, st));
        if (ok && mm.has_hi) run("qhigh_split", launch_qhigh(me, se, false, st));
        if (ok) run("pair_energy", launch_pair_energy(me, se, epv, st));
    }
    if (ok) run("atom_ev", launch_atom_ev(se, epv, evat, st));
    return ok;
}

void collect_profile(mlip_ctx* c) {
    for (size_t k = 0; k < c->ev_used; ++k) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->events[k].a, c->events[k].b);
        const std::string name = c->events[k].name;
        size_t j = 0;
        while (j < c->prof_names.size() && c->prof_names[j] != name) ++j;
        if (j == c->prof_names.size()) {
            c->prof_names.push_back(name);
            c->prof_ms.push_back(0.0);
            c->prof_count.push_back(0);
        }
        c->prof_ms[j] += ms;
        c->prof_count[j] += 1;
    }
    c->ev_used = 0;
}

}  // namespace mlip

namespace {

int lammps_compute(mlip_ctx* c, int nlocal, int nghost, const double* x, const int* type, const int64_t* tag,
                   int inum, const int* ilist, const int* numneigh, const int* const* firstneigh,
                   const int* neigh_off, const int* neigh_idx, int neigh_full, int eflag, int vflag, double* f,
                   double* eng, double* virial, double* eatom, double* vatom) {
    (void)eflag;
    (void)vflag;
    LammpsLists L;
    const int rc = lammps_lists(c, nlocal, nghost, x, type, tag, inum, ilist, numneigh, firstneigh, neigh_off,
                                neigh_idx, neigh_full, L);
    if (rc != MLIP_OK) return rc;
    c->resident = false;
    c->nstruct = 0;
    if (!setup_system(c, L.nall, inum, L.npairs, L.x, L.elem, L.elem_own, L.poff, L.pnbr, &L.pcen, nullptr, L.gof,
                      L.gen))
        return MLIP_ERR_DEVICE;
    if (!run_pipeline(c)) return MLIP_ERR_DEVICE;
    double* epv = nullptr;
    double* evat = nullptr;
    if ((eatom || vatom) && c->atom_mode == MLIP_ATOM_PAIR_SPLIT) {
        if (!c->epv.ensure(7 * static_cast<size_t>(std::max<int64_t>(L.npairs, 1))) ||
            !c->evat.ensure(7 * static_cast<size_t>(std::max(L.nall, 1)))) {
            c->err = "device allocation failed for per-atom energies";
            return MLIP_ERR_DEVICE;
        }
        epv = c->epv.p;
        evat = c->evat.p;
    }
    return lammps_results(c, L.nall, inum, ilist, f, eng, virial, eatom, vatom, epv, evat);
}

}  // namespace

namespace mlip {

int lammps_lists(mlip_ctx* c, int nlocal, int nghost, const double* x, const int* type, const int64_t* tag, int inum,
                 const int* ilist, const int* numneigh, const int* const* firstneigh, const int* neigh_off,
                 const int* neigh_idx, int neigh_full, LammpsLists& L) {
    const int nall = nlocal + nghost;
    if (nlocal < 0 || nghost < 0 || inum < 0 || inum > nlocal || !x || !type || (inum > 0 && (!ilist || !numneigh))) {
        c->err = "bad atom counts or null arrays";
        return MLIP_ERR_ARG;
    }
    L.nall = nall;
    L.elem.assign(nall, 0);
    for (int k = 0; k < nall; ++k) {
        const int t = type[k];
        if (t < 1 || t > static_cast<int>(c->elem_of_type.size()) || c->elem_of_type[t - 1] < 0) {
            c->err = "atom type without an element mapping";
            return MLIP_ERR_ARG;
        }
        L.elem[k] = c->elem_of_type[t - 1];
    }
    // owners: ilist order; half lists need the owner of each ghost (by tag)
    std::vector<int> owner_of_local(nlocal, -1);
    for (int ii = 0; ii < inum; ++ii) {
        if (ilist[ii] < 0 || ilist[ii] >= nlocal) {
            c->err = "ilist entry out of range";
            return MLIP_ERR_ARG;
        }
        owner_of_local[ilist[ii]] = ii;
    }
    std::unordered_map<int64_t, int> local_of_tag;
    if (!neigh_full) {
        if (!tag) {
            c->err = "half neighbor lists need atom tags (newton on, single rank)";
            return MLIP_ERR_ARG;
        }
        local_of_tag.reserve(nlocal * 2);
        for (int i = 0; i < nlocal; ++i) local_of_tag[tag[i]] = i;
    }
    auto list_of = [&](int ii, int& n) -> const int* {
        const int i = ilist[ii];
        n = numneigh[i];
        return firstneigh ? firstneigh[i] : neigh_idx + neigh_off[ii];
    };
    std::vector<std::vector<std::pair<int32_t, int32_t>>> per(inum);  // (centre pos, neighbor pos)
    for (int ii = 0; ii < inum; ++ii) {
        int n = 0;
        const int* js = list_of(ii, n);
        const int i = ilist[ii];
        if (n < 0 || (n > 0 && !js)) {
            c->err = "bad neighbor count or list";
            return MLIP_ERR_ARG;
        }
        for (int jj = 0; jj < n; ++jj) {
            const int j = js[jj];
            if (j < 0 || j >= nall) {
                c->err = "neighbor index out of range";
                return MLIP_ERR_ARG;
            }
            per[ii].push_back({i, j});
            if (!neigh_full) {
                auto it = local_of_tag.find(tag[j]);
                if (it == local_of_tag.end() || owner_of_local[it->second] < 0) {
                    c->err = "neighbor tag without a local owner (half lists are single-rank)";
                    return MLIP_ERR_ARG;
                }
                per[owner_of_local[it->second]].push_back({j, i});
            }
        }
    }
    L.poff.assign(inum + 1, 0);
    L.pnbr.clear();
    L.pcen.clear();
    for (int ii = 0; ii < inum; ++ii) {
        L.poff[ii + 1] = L.poff[ii] + static_cast<int32_t>(per[ii].size());
        for (const auto& e : per[ii]) {
            L.pcen.push_back(e.first);
            L.pnbr.push_back(e.second);
        }
    }
    L.npairs = static_cast<int64_t>(L.pnbr.size());
    if (L.npairs > (INT32_MAX - 1) / 2) {
        c->err = "more than 2^30 pairs";  // gather offsets reach 2 * npairs
        return MLIP_ERR_ARG;
    }
    // owners are numbered by ilist position; centre and neighbor positions
    // keep their LAMMPS indices through pcen / pnbr
    L.x.assign(x, x + 3 * static_cast<size_t>(nall));
    L.elem_own.assign(inum, 0);
    for (int ii = 0; ii < inum; ++ii) L.elem_own[ii] = L.elem[ilist[ii]];
    build_gather(nall, L.poff, L.pnbr, &L.pcen, L.gof, L.gen);
    return MLIP_OK;
}

int lammps_results(mlip_ctx* c, int nall, int inum, const int* ilist, double* f, double* eng, double* virial,
                   double* eatom, double* vatom, double* epv, double* evat) {
    std::vector<double> fh(3 * static_cast<size_t>(nall)), outv(7);
    if (!hip_ok(c, hipMemcpyAsync(fh.data(), c->ds.f, fh.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                "copy forces") ||
        !hip_ok(c, hipMemcpyAsync(outv.data(), c->ds.out, 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                "copy energy") ||
        !hip_ok(c, hipStreamSynchronize(c->stream), "compute"))
        return MLIP_ERR_DEVICE;
    collect_profile(c);
    if (f)
        for (size_t k = 0; k < fh.size(); ++k) f[k] += fh[k];
    if (eng) *eng += outv[0];
    if (virial)
        for (int k = 0; k < 6; ++k) virial[k] += outv[1 + k];
    if ((eatom || vatom) && c->atom_mode == MLIP_ATOM_PAIR_SPLIT) {
        // the reference's ev_tally_xyz split: every position gets half of
        // each pair energy and pair virial it takes part in
        if (!run_atom_split(c, epv, evat)) return MLIP_ERR_DEVICE;
        std::vector<double> ev(7 * static_cast<size_t>(nall));
        if (!hip_ok(c, hipMemcpyAsync(ev.data(), evat, ev.size() * sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream),
                    "copy per-atom energy") ||
            !hip_ok(c, hipStreamSynchronize(c->stream), "per-atom energy"))
            return MLIP_ERR_DEVICE;
        for (int k = 0; k < nall; ++k) {
            if (eatom) eatom[k] += ev[7 * static_cast<size_t>(k)];
            if (vatom)
                for (int q = 0; q < 6; ++q)
                    vatom[6 * static_cast<size_t>(k) + q] += ev[7 * static_cast<size_t>(k) + 1 + q];
        }
    } else if (eatom || vatom) {
        // site energies E_i and the virial of the pairs each centre owns
        const size_t nslot = static_cast<size_t>(c->ds.nblocks) * kBlock;
        std::vector<double> es(nslot), vo(static_cast<size_t>(inum) * 6);
        std::vector<int32_t> sl(inum);
        if (!hip_ok(c, hipMemcpyAsync(es.data(), c->ds.e_slot, nslot * sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream),
                    "copy site energies") ||
            !hip_ok(c, hipMemcpyAsync(vo.data(), c->ds.vown, vo.size() * sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream),
                    "copy site virials") ||
            !hip_ok(c, hipMemcpyAsync(sl.data(), c->ds.slot_of_atom, inum * sizeof(int32_t), hipMemcpyDeviceToHost,
                                      c->stream),
                    "copy slots") ||
            !hip_ok(c, hipStreamSynchronize(c->stream), "site energies"))
            return MLIP_ERR_DEVICE;
        for (int ii = 0; ii < inum; ++ii) {
            if (eatom) eatom[ilist[ii]] += es[sl[ii]];
            if (vatom)
                for (int k = 0; k < 6; ++k) vatom[6 * ilist[ii] + k] += vo[6 * ii + k];
        }
    }
    return MLIP_OK;
}

}  // namespace mlip

namespace {

}  // namespace

extern "C" {

int mlip_settings(int narg) { return narg == 0 ? MLIP_OK : MLIP_ERR_ARG; }

int mlip_create(const char* potfile, int ntypes, const char* const* elem_of_type, mlip_ctx** out, char* errbuf,
                size_t errlen) {
    if (!out || !potfile || ntypes < 1) {
        set_err(errbuf, errlen, "Incorrect args for pair coefficients");
        return MLIP_ERR_ARG;
    }
    *out = nullptr;
    GtinvTable table;
    std::string e;
    if (!load_builtin_gtinv(table, e)) {
        set_err(errbuf, errlen, e);
        return MLIP_ERR_MODEL;
    }
    auto* c = new mlip_ctx();
    if (!load_model(potfile, table, c->model, e)) {
        set_err(errbuf, errlen, e);
        delete c;
        return e.rfind("cannot open", 0) == 0 ? MLIP_ERR_FILE : MLIP_ERR_MODEL;
    }
    c->elem_of_type.assign(ntypes, -1);
    for (int t = 0; t < ntypes; ++t) {
        const char* name = elem_of_type ? elem_of_type[t] : nullptr;
        if (!name || std::strcmp(name, "NULL") == 0) continue;
        const auto& el = c->model.elements;
        auto it = std::find(el.begin(), el.end(), std::string(name));
        if (it == el.end()) {
            set_err(errbuf, errlen, std::string("element ") + name + " not in potential file");
            delete c;
            return MLIP_ERR_ARG;
        }
        c->elem_of_type[t] = static_cast<int>(it - el.begin());
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_err(errbuf, errlen, "no HIP device available: the gtinv path runs on the GPU only");
        delete c;
        return MLIP_ERR_DEVICE;
    }
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err(errbuf, errlen, "failed to create a HIP stream");
        mlip_destroy(c);
        return MLIP_ERR_DEVICE;
    }
    if (!upload_model(c)) {
        set_err(errbuf, errlen, c->err);
        mlip_destroy(c);
        return MLIP_ERR_DEVICE;
    }
    *out = c;
    return MLIP_OK;
}

void mlip_destroy(mlip_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    DBuf<double>* dbs[] = {&c->d_coef, &c->d_weight, &c->d_weight_e, &c->d_wmat, &c->d_linp, &c->d_c3w, &c->x,
                           &c->pshift, &c->epv, &c->evat,
                           &c->A, &c->G, &c->D, &c->Hq, &c->H, &c->Ep, &c->Eq12, &c->Eq34, &c->gpair,
                           &c->vown, &c->f, &c->e_slot, &c->red, &c->out, &c->sout};
    for (auto* b : dbs) b->release();
    DBuf<int32_t>* ibs[] = {&c->d_q12off, &c->d_c3off, &c->d_c3bc, &c->d_pdeg, &c->elem, &c->poff, &c->pnbr, &c->pcen,
                            &c->atom_of_slot, &c->slot_of_atom, &c->block_type, &c->goff, &c->gent,
                            &c->soff, &c->d_zlist};
    for (auto* b : ibs) b->release();
    c->d_q12.release();
    c->d_fq12.release();
    c->d_vadj.release();
    c->d_vfeat.release();
    c->d_gadj.release();
    c->d_gfeat.release();
    c->d_yt64.release();
    c->d_yt32.release();
    c->d_pairw.release();
    c->d_phoff.release();
    c->d_phent.release();
    c->d_hvis.release();
    c->d_hterm.release();
    c->d_hcoef.release();
    drop_graph(c);
    pool_free(c);
    nb_free(c);
    train_free(c);
    c->soff.release();
    c->sout.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

double mlip_cutoff(const mlip_ctx* c) { return c ? c->model.cutoff : 0.0; }

int mlip_masses(const mlip_ctx* c, double* mass) {
    if (!c || !mass) return MLIP_ERR_ARG;
    for (size_t t = 0; t < c->elem_of_type.size(); ++t)
        mass[t] = c->elem_of_type[t] >= 0 ? c->model.masses[c->elem_of_type[t]] : 0.0;
    return MLIP_OK;
}

int mlip_set_precision(mlip_ctx* c, int mode) {
    if (!c || (mode != MLIP_FP64 && mode != MLIP_MIXED)) return MLIP_ERR_ARG;
    if (c->precision != mode) drop_graph(c);
    c->precision = mode;
    return MLIP_OK;
}

int mlip_set_atom_mode(mlip_ctx* c, int mode) {
    if (!c || (mode != MLIP_ATOM_PAIR_SPLIT && mode != MLIP_ATOM_SITE)) return MLIP_ERR_ARG;
    c->atom_mode = mode;
    return MLIP_OK;
}

const char* mlip_last_error(const mlip_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mlip_compute(mlip_ctx* c, int nlocal, int nghost, const double* x, const int* type, const int64_t* tag, int inum,
                 const int* ilist, const int* numneigh, const int* const* firstneigh, int neigh_full, int eflag,
                 int vflag, double* f, double* eng, double* virial, double* eatom, double* vatom) {
    if (!c || !firstneigh || !f) return MLIP_ERR_ARG;
    return lammps_compute(c, nlocal, nghost, x, type, tag, inum, ilist, numneigh, firstneigh, nullptr, nullptr,
                          neigh_full, eflag, vflag, f, eng, virial, eatom, vatom);
}

int mlip_compute_flat(mlip_ctx* c, int nlocal, int nghost, const double* x, const int* type, const int64_t* tag,
                      int inum, const int* ilist, const int* numneigh, const int* neigh_off, const int* neigh_idx,
                      int neigh_full, int eflag, int vflag, double* f, double* eng, double* virial, double* eatom,
                      double* vatom) {
    if (!c || !neigh_off || !neigh_idx || !f) return MLIP_ERR_ARG;
    return lammps_compute(c, nlocal, nghost, x, type, tag, inum, ilist, numneigh, nullptr, neigh_off, neigh_idx,
                          neigh_full, eflag, vflag, f, eng, virial, eatom, vatom);
}

int mlip_resident_set(mlip_ctx* c, int natoms, const double* x, const int* type, int64_t npairs,
                      const int64_t* nbr_off, const int* nbr_idx, const double* nbr_shift) {
    if (!c || natoms < 0 || !x || !type || !nbr_off || (npairs > 0 && !nbr_idx)) return MLIP_ERR_ARG;
    if (npairs > (INT32_MAX - 1) / 2) {
        c->err = "more than 2^30 pairs";  // gather offsets reach 2 * npairs
        return MLIP_ERR_ARG;
    }
    std::vector<int32_t> elem(natoms), poff(natoms + 1), pnbr(nbr_idx, nbr_idx + npairs);
    for (int i = 0; i < natoms; ++i) {
        const int t = type[i];
        if (t < 1 || t > static_cast<int>(c->elem_of_type.size()) || c->elem_of_type[t - 1] < 0) {
            c->err = "atom type without an element mapping";
            return MLIP_ERR_ARG;
        }
        elem[i] = c->elem_of_type[t - 1];
    }
    for (int i = 0; i <= natoms; ++i) poff[i] = static_cast<int32_t>(nbr_off[i]);
    for (int64_t p = 0; p < npairs; ++p)
        if (pnbr[p] < 0 || pnbr[p] >= natoms) {
            c->err = "neighbor index out of range";
            return MLIP_ERR_ARG;
        }
    std::vector<int32_t> gof, gen;
    build_gather(natoms, poff, pnbr, nullptr, gof, gen);
    std::vector<double> xv(x, x + 3 * static_cast<size_t>(natoms));
    std::vector<double> sh;
    if (nbr_shift) sh.assign(nbr_shift, nbr_shift + 3 * npairs);
    c->nstruct = 0;
    if (!setup_system(c, natoms, natoms, npairs, xv, elem, elem, poff, pnbr, nullptr, nbr_shift ? &sh : nullptr, gof,
                      gen))
        return MLIP_ERR_DEVICE;
    if (!hip_ok(c, hipStreamSynchronize(c->stream), "resident upload")) return MLIP_ERR_DEVICE;
    c->resident = true;
    c->resident_natoms = natoms;
    return MLIP_OK;
}

int mlip_resident_update_positions(mlip_ctx* c, const double* x) {
    if (!c || !x) return MLIP_ERR_ARG;
    if (!c->resident) {
        c->err = "no resident structure";
        return MLIP_ERR_STATE;
    }
    return hip_ok(c, hipMemcpyAsync(c->x.p, x, 3 * sizeof(double) * c->resident_natoms, hipMemcpyHostToDevice,
                                    c->stream),
                  "update positions")
               ? MLIP_OK
               : MLIP_ERR_DEVICE;
}

int mlip_resident_compute(mlip_ctx* c) {
    if (!c) return MLIP_ERR_ARG;
    if (!c->resident) {
        c->err = "no resident structure";
        return MLIP_ERR_STATE;
    }
    return run_eval(c) ? MLIP_OK : MLIP_ERR_DEVICE;
}

int mlip_resident_sync(mlip_ctx* c) {
    if (!c) return MLIP_ERR_ARG;
    if (!hip_ok(c, hipStreamSynchronize(c->stream), "resident compute")) return MLIP_ERR_DEVICE;
    collect_profile(c);
    return MLIP_OK;
}

int mlip_resident_get(mlip_ctx* c, double* f, double* energy, double* virial, double* eatom) {
    if (!c) return MLIP_ERR_ARG;
    if (!c->resident) {
        c->err = "no resident structure";
        return MLIP_ERR_STATE;
    }
    const int n = c->resident_natoms;
    double outv[7];
    if (f && !hip_ok(c, hipMemcpyAsync(f, c->f.p, 3 * sizeof(double) * n, hipMemcpyDeviceToHost, c->stream), "forces"))
        return MLIP_ERR_DEVICE;
    if (!hip_ok(c, hipMemcpyAsync(outv, c->out.p, sizeof outv, hipMemcpyDeviceToHost, c->stream), "energy") ||
        !hip_ok(c, hipStreamSynchronize(c->stream), "readback"))
        return MLIP_ERR_DEVICE;
    if (energy) *energy = outv[0];
    if (virial)
        for (int k = 0; k < 6; ++k) virial[k] = outv[1 + k];
    if (eatom) {
        const size_t nslot = static_cast<size_t>(c->ds.nblocks) * kBlock;
        std::vector<double> es(nslot);
        std::vector<int32_t> sl(n);
        (void)hipMemcpy(es.data(), c->e_slot.p, nslot * sizeof(double), hipMemcpyDeviceToHost);
        (void)hipMemcpy(sl.data(), c->slot_of_atom.p, n * sizeof(int32_t), hipMemcpyDeviceToHost);
        for (int i = 0; i < n; ++i) eatom[i] = es[sl[i]];
    }
    return MLIP_OK;
}

int mlip_batch_set(mlip_ctx* c, int nstruct, const int* natoms, const double* cells, const double* x,
                   const int* type) {
    if (!c || nstruct <= 0 || !natoms || !cells || !x || !type) return MLIP_ERR_ARG;
    std::vector<int32_t> soff(nstruct + 1, 0);
    for (int k = 0; k < nstruct; ++k) {
        if (natoms[k] < 0) return MLIP_ERR_ARG;
        soff[k + 1] = soff[k] + natoms[k];
    }
    const int n = soff[nstruct];
    std::vector<int32_t> elem(n), poff(n + 1, 0), pnbr;
    std::vector<double> sh;
    for (int i = 0; i < n; ++i) {
        const int t = type[i];
        if (t < 1 || t > static_cast<int>(c->elem_of_type.size()) || c->elem_of_type[t - 1] < 0) {
            c->err = "atom type without an element mapping";
            return MLIP_ERR_ARG;
        }
        elem[i] = c->elem_of_type[t - 1];
    }
    c->nstruct = 0;
    if (c->nb_device) {
        const int rc = setup_cells_device(c, nstruct, natoms, cells, x, elem);
        return rc != MLIP_OK ? rc : finish_batch(c, nstruct, soff, n);
    }
    // per-structure neighbor lists on host threads, concatenated in order
    std::vector<NeighborList> lists(nstruct);
    std::vector<std::string> errs(nstruct);
    std::vector<char> bad(nstruct, 0);
    {
        std::atomic<int> next{0};
        auto work = [&]() {
            for (int k = next++; k < nstruct; k = next++)
                bad[k] = !build_full_list(cells + 9 * k, natoms[k], x + 3 * static_cast<size_t>(soff[k]),
                                          c->model.cutoff, lists[k], errs[k]);
        };
        const int nt = std::max(1, std::min<int>(nstruct, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    }
    for (int k = 0; k < nstruct; ++k) {
        if (bad[k]) {
            c->err = "structure " + std::to_string(k) + ": " + errs[k];
            return MLIP_ERR_ARG;
        }
        NeighborList nl = std::move(lists[k]);
        for (int i = 0; i < natoms[k]; ++i) poff[soff[k] + i + 1] = static_cast<int32_t>(pnbr.size() + nl.off[i + 1]);
        for (int32_t j : nl.idx) pnbr.push_back(soff[k] + j);
        sh.insert(sh.end(), nl.shift.begin(), nl.shift.end());
        if (pnbr.size() > static_cast<size_t>((INT32_MAX - 1) / 2)) {
            c->err = "more than 2^30 pairs in the batch";
            return MLIP_ERR_ARG;
        }
    }
    std::vector<int32_t> gof, gen;
    build_gather(n, poff, pnbr, nullptr, gof, gen);
    std::vector<double> xv(x, x + 3 * static_cast<size_t>(n));
    if (!setup_system(c, n, n, static_cast<int64_t>(pnbr.size()), xv, elem, elem, poff, pnbr, nullptr, &sh, gof, gen))
        return MLIP_ERR_DEVICE;
    return finish_batch(c, nstruct, soff, n);
}

int mlip_train_features(mlip_ctx* c, double* X, double* P, double* FF, double* SF) {
    if (!c) return MLIP_ERR_ARG;
    return train_features(c, X, P, FF, SF);
}

int mlip_set_neighbor_builder(mlip_ctx* c, int device) {
    if (!c || (device != 0 && device != 1)) return MLIP_ERR_ARG;
    c->nb_device = device == 1;
    return MLIP_OK;
}

int mlip_resident_set_cell(mlip_ctx* c, int natoms, const double* cell9, const double* x, const int* type) {
    if (!c || natoms < 0 || !cell9 || (natoms > 0 && (!x || !type))) return MLIP_ERR_ARG;
    std::vector<int32_t> elem(natoms);
    for (int i = 0; i < natoms; ++i) {
        const int t = type[i];
        if (t < 1 || t > static_cast<int>(c->elem_of_type.size()) || c->elem_of_type[t - 1] < 0) {
            c->err = "atom type without an element mapping";
            return MLIP_ERR_ARG;
        }
        elem[i] = c->elem_of_type[t - 1];
    }
    c->resident = false;
    c->nstruct = 0;
    const int32_t na = natoms;
    const int rc = setup_cells_device(c, 1, &na, cell9, x, elem);
    if (rc != MLIP_OK) return rc;
    if (!hip_ok(c, hipStreamSynchronize(c->stream), "resident build")) return MLIP_ERR_DEVICE;
    c->resident = true;
    c->resident_natoms = natoms;
    return MLIP_OK;
}

int64_t mlip_resident_npairs(const mlip_ctx* c) { return c && c->resident ? c->npairs : -1; }

int mlip_resident_lists(mlip_ctx* c, int64_t* off, int* idx, double* shift, int* goff, int* gent) {
    if (!c) return MLIP_ERR_ARG;
    if (!c->resident) {
        c->err = "no resident structure";
        return MLIP_ERR_STATE;
    }
    const size_t n = static_cast<size_t>(c->resident_natoms), np = static_cast<size_t>(c->npairs);
    hipStream_t st = c->stream;
    std::vector<int32_t> o32(n + 1);
    bool ok = hip_ok(c, hipMemcpyAsync(o32.data(), c->poff.p, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st),
                     "pair offsets");
    if (ok && idx && np)
        ok = hip_ok(c, hipMemcpyAsync(idx, c->pnbr.p, np * sizeof(int32_t), hipMemcpyDeviceToHost, st), "pairs");
    if (ok && shift && np) {
        if (!c->ds.pshift) {
            c->err = "the resident lists carry no image shifts";
            return MLIP_ERR_STATE;
        }
        ok = hip_ok(c, hipMemcpyAsync(shift, c->pshift.p, 3 * np * sizeof(double), hipMemcpyDeviceToHost, st),
                    "shifts");
    }
    if (ok && goff)
        ok = hip_ok(c, hipMemcpyAsync(goff, c->goff.p, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st),
                    "gather offsets");
    if (ok && gent && np)
        ok = hip_ok(c, hipMemcpyAsync(gent, c->gent.p, 2 * np * sizeof(int32_t), hipMemcpyDeviceToHost, st),
                    "gather entries");
    if (!ok || !hip_ok(c, hipStreamSynchronize(st), "list readback")) return MLIP_ERR_DEVICE;
    if (off)
        for (size_t i = 0; i <= n; ++i) off[i] = o32[i];
    return MLIP_OK;
}

int mlip_random_structures(int nstruct, int nmin, int nmax, double density, double rmin, uint64_t seed,
                           int* natoms, double* cells, double* x, int* type, int64_t cap_atoms) {
    if (nstruct < 0 || nmin < 1 || nmax < nmin || !(density > 0) || !natoms) return MLIP_ERR_ARG;
    int64_t total = 0;
    for (int k = 0; k < nstruct; ++k) {
        natoms[k] = random_structure_size(seed, k, nmin, nmax);
        total += natoms[k];
    }
    if (!x) return MLIP_OK;  // sizes only
    if (!cells || !type || total > cap_atoms) return MLIP_ERR_ARG;
    int64_t at = 0;
    for (int k = 0; k < nstruct; ++k) {
        if (!random_structure(seed, k, natoms[k], density, rmin, cells + 9 * k, x + 3 * at, type + at))
            return MLIP_ERR_ARG;
        at += natoms[k];
    }
    return MLIP_OK;
}

int mlip_batch_get(mlip_ctx* c, double* energies, double* virials, double* f) {
    if (!c) return MLIP_ERR_ARG;
    if (c->nstruct <= 0) {
        c->err = "no batch";
        return MLIP_ERR_STATE;
    }
    std::vector<double> o(7 * static_cast<size_t>(c->nstruct));
    if (!hip_ok(c, hipMemcpyAsync(o.data(), c->sout.p, o.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                "batch results"))
        return MLIP_ERR_DEVICE;
    if (f && !hip_ok(c, hipMemcpyAsync(f, c->f.p, 3 * sizeof(double) * c->resident_natoms, hipMemcpyDeviceToHost,
                                       c->stream),
                     "batch forces"))
        return MLIP_ERR_DEVICE;
    if (!hip_ok(c, hipStreamSynchronize(c->stream), "batch readback")) return MLIP_ERR_DEVICE;
    for (int k = 0; k < c->nstruct; ++k) {
        if (energies) energies[k] = o[7 * k];
        if (virials)
            for (int q = 0; q < 6; ++q) virials[6 * k + q] = o[7 * k + 1 + q];
    }
    return MLIP_OK;
}

int mlip_graph_mode(mlip_ctx* c, int on) {
    if (!c) return MLIP_ERR_ARG;
    if ((on != 0) != c->use_graph) drop_graph(c);
    c->use_graph = on != 0;
    return MLIP_OK;
}

int mlip_autotune(mlip_ctx* c, int reps, int* chosen) {
    if (!c || reps < 1) return MLIP_ERR_ARG;
    if (!c->resident) {
        c->err = "autotune needs a resident structure or batch";
        return MLIP_ERR_STATE;
    }
    const int cand[4] = {64, 128, 192, 256};
    const bool prof = c->prof;
    double best = 1e300;
    int best_nt = c->force_threads;
    for (int nt : cand) {
        c->force_threads = nt;
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return MLIP_ERR_DEVICE;
        (void)hipEventRecord(a, c->stream);
        hipError_t le = hipSuccess;
        for (int r = 0; r < reps && le == hipSuccess; ++r) le = launch_force(c->dm, c->ds, c->precision, nt, c->stream);
        (void)hipEventRecord(b, c->stream);
        float ms = 0.0f;
        const bool ok = hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (le != hipSuccess) {
            c->err = std::string("force kernel launch: ") + hipGetErrorString(le);
            return MLIP_ERR_DEVICE;
        }
        if (!ok) return MLIP_ERR_DEVICE;
        if (ms < best) {
            best = ms;
            best_nt = nt;
        }
    }
    c->force_threads = best_nt;
    c->prof = prof;
    // a_nlm: the matrix-core kernel against the register one, where the
    // model fits the former
    if (!c->model.pair_model && anlm_has_mfma(c->dm)) {
        double tk[2] = {0.0, 0.0};
        for (int kind = 0; kind < 2; ++kind) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return MLIP_ERR_DEVICE;
            // each with what it leaves to the order-1/2 feature pass
            const bool ff = c->dm.n_p > 0 && anlm_forms_features(c->dm, c->precision, kind);
            auto once = [&] {
                hipError_t e = launch_anlm(c->dm, c->ds, c->precision, c->stream, kind, ff);
                if (e == hipSuccess && c->dm.n_p > 0 && !ff) e = launch_q12(c->dm, c->ds, true, c->stream);
                return e;
            };
            hipError_t le = once();
            (void)hipEventRecord(a, c->stream);
            for (int r = 0; r < reps && le == hipSuccess; ++r) le = once();
            (void)hipEventRecord(b, c->stream);
            float ms = 0.0f;
            const bool ok = le == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
                            hipEventElapsedTime(&ms, a, b) == hipSuccess;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            if (!ok) {
                c->err = std::string("a_nlm kernel timing: ") + hipGetErrorString(le);
                return MLIP_ERR_DEVICE;
            }
            tk[kind] = ms;
        }
        c->anlm_kind = tk[kAnlmMfma] <= tk[kAnlmValu] ? kAnlmMfma : kAnlmValu;
    }
    drop_graph(c);
    if (chosen) *chosen = best_nt;
    return MLIP_OK;
}

int mlip_set_anlm_kernel(mlip_ctx* c, int kind) {
    if (!c || (kind != kAnlmValu && kind != kAnlmMfma)) return MLIP_ERR_ARG;
    if (kind != c->anlm_kind) drop_graph(c);
    c->anlm_kind = kind;
    return MLIP_OK;
}

int mlip_get_anlm_kernel(const mlip_ctx* c) { return c ? c->anlm_kind : -1; }

int mlip_autotune_q34(mlip_ctx* c, int reps, int* plan2) {
    if (!c || reps < 1) return MLIP_ERR_ARG;
    if (!c->resident) {
        c->err = "autotune needs a resident structure or batch";
        return MLIP_ERR_STATE;
    }
    Model& m = c->model;
    if (!m.has_order34 || m.pair_model) {
        if (plan2) plan2[0] = m.plan_region, plan2[1] = m.plan_visit, plan2[2] = m.plan_pair;
        return MLIP_OK;
    }
    // the work models tried (terms per scheduling region, per visit); the
    // structure's buffers hold the last evaluation's a_nlm and H, and the
    // timed launches only rewrite G and the order-3/4 energies, which the
    // next evaluation writes again
    // work models (terms per scheduling region, per visit) times the largest
    // L of species-paired rows (pairing forms fewer products but leaves
    // fewer, larger parts to balance over the waves)
    const int work[][2] = {{3, 24}, {0, 0}, {2, 8}, {3, 12}, {3, 48}, {6, 24}, {1, 24}, {4, 32}};
    const int pairs[] = {kPairMaxL, kPairMaxL - 1, kPairMaxL - 2};
    int cand[24][3];
    int ncand = 0;
    for (int pl : pairs)
        for (const auto& w : work) {
            cand[ncand][0] = w[0];
            cand[ncand][1] = w[1];
            cand[ncand][2] = std::max(pl, -1);
            ++ncand;
        }
    double best = 1e300;
    int bi = 0;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return MLIP_ERR_DEVICE;
    int rc = MLIP_OK;
    for (int k = 0; k < ncand && rc == MLIP_OK; ++k) {
        plan_q34(m, cand[k][0], cand[k][1], cand[k][2]);
        if (!upload_q34_plan(c) || !hip_ok(c, hipStreamSynchronize(c->stream), "plan upload")) {
            rc = MLIP_ERR_DEVICE;
            break;
        }
        hipError_t le = launch_q34_adj(c->dm, c->ds, c->stream);  // warm-up
        (void)hipEventRecord(a, c->stream);
        for (int r = 0; r < reps && le == hipSuccess; ++r) le = launch_q34_adj(c->dm, c->ds, c->stream);
        (void)hipEventRecord(b, c->stream);
        float ms = 0.0f;
        if (le != hipSuccess || hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
            c->err = std::string("order-3/4 plan timing: ") + hipGetErrorString(le);
            rc = MLIP_ERR_DEVICE;
            break;
        }
        if (ms < best) {
            best = ms;
            bi = k;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    plan_q34(m, cand[bi][0], cand[bi][1], cand[bi][2]);
    if (!upload_q34_plan(c) || !hip_ok(c, hipStreamSynchronize(c->stream), "plan upload")) return MLIP_ERR_DEVICE;
    drop_graph(c);
    if (plan2) plan2[0] = cand[bi][0], plan2[1] = cand[bi][1], plan2[2] = cand[bi][2];
    return rc;
}

int mlip_set_q34_plan(mlip_ctx* c, int region_cost, int visit_cost, int pair_lmax) {
    if (!c || region_cost < 0 || visit_cost < 0 || pair_lmax < -1 || region_cost > 1000 || visit_cost > 10000) {
        if (c) c->err = "q34 plan: work-model constants must be >= 0, the pairing limit >= -1";
        return MLIP_ERR_ARG;
    }
    Model& m = c->model;
    if (!m.has_order34 || m.pair_model) return MLIP_OK;
    plan_q34(m, region_cost, visit_cost, pair_lmax);
    if (!upload_q34_plan(c) || !hip_ok(c, hipStreamSynchronize(c->stream), "plan upload")) return MLIP_ERR_DEVICE;
    drop_graph(c);
    return MLIP_OK;
}

int mlip_set_force_threads(mlip_ctx* c, int threads) {
    if (!c) return MLIP_ERR_ARG;
    if (threads != 64 && threads != 128 && threads != 192 && threads != 256) {
        c->err = "force threads must be 64, 128, 192 or 256";
        return MLIP_ERR_ARG;
    }
    if (threads != c->force_threads) drop_graph(c);
    c->force_threads = threads;
    return MLIP_OK;
}

int mlip_profile_enable(mlip_ctx* c, int on) {
    if (!c) return MLIP_ERR_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    collect_profile(c);
    c->prof = on != 0;
    c->prof_names.clear();
    c->prof_ms.clear();
    c->prof_count.clear();
    return MLIP_OK;
}

int mlip_profile_read(mlip_ctx* c, const char** names, double* ms, int* count, int cap) {
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    collect_profile(c);
    int n = 0;
    for (size_t k = 0; k < c->prof_names.size() && n < cap; ++k, ++n) {
        names[n] = c->prof_names[k].c_str();
        ms[n] = c->prof_ms[k];
        if (count) count[n] = c->prof_count[k];
    }
    return n;
}

// Algorithmic FP64 operation counts of one evaluation of the current
// structure, per kernel (adds and multiplies count 1, fused multiply-adds 2):
//   [0] a_nlm       per pair 4 n_fn n_lmh (accumulate) + 8 n_lmh (Y_lm) + 8 n_fn
//   [1] order 1/2   per centre and n: 4 per order-1 family, 20 (L+1) per order-2 family
//   [2] order 3/4 polynomial features  per valid term 11, + 6 per (m_a, m_b) at order 4
//   [3] order 3/4 adjoint              per valid term 10, + 6 per (m_a, m_b) at order 4,
//                                      + 2 per m_a, + 10 (L+1) per group (energy by Euler)
//   [4] polynomial  2 w^2 (quadratic form) + 6 w + 4 per cubic entry, per centre, w = the
//                   feature columns its species sees
//   [5] forces      per pair 8 n_fn n_lmh (contraction) + 21 n_lmh (recursion and
//                   the three column sums) + 60 (maxl + 1) (per-column fold) + 14 n_fn
//   [6] total
int mlip_flops(const mlip_ctx* c, double* out7) {
    if (!c || !out7) return MLIP_ERR_ARG;
    const Model& m = c->model;
    for (int k = 0; k < 7; ++k) out7[k] = 0.0;
    const double pairs = static_cast<double>(c->npairs);
    out7[0] = pairs * (4.0 * m.n_fn * m.n_lmh + 8.0 * m.n_lmh + 8.0 * m.n_fn);
    out7[5] = pairs * (8.0 * m.n_fn * m.n_lmh + 21.0 * m.n_lmh + 60.0 * (m.maxl + 1) + 14.0 * m.n_fn);
    for (int t = 0; t < m.n_type && t < static_cast<int>(c->owners_per_type.size()); ++t) {
        const double na = static_cast<double>(c->owners_per_type[t]) * m.n_fn;
        const TypeTables& tt = m.types[t];
        // orders 1/2: feature pass 6 (L + 1) per polynomial family (2 for
        // order 1), adjoint 14 (L + 1) (2); adjoint rows L <= q12_fuse are
        // counted with the order-3/4 kernel that runs them
        double q12 = 0.0;
        for (const VisitQ12& v : tt.visits_q12) {
            int L = 0;
            for (int l = 0; l <= m.maxl; ++l)
                if (&v - tt.visits_q12.data() >= tt.q12_off[l]) L = l;
            if (v.pslot >= 0) q12 += v.order == 1 ? 2.0 : 6.0 * (L + 1);
            if (!q12_fused(c->dm, L)) q12 += v.order == 1 ? 2.0 : 14.0 * (L + 1);
        }
        out7[1] += na * q12;
        auto terms = [&](const std::vector<GroupQ34>& gs, const std::vector<VisitQ34>& vs, bool adj) {
            double f = 0.0;
            for (const GroupQ34& g : gs)
                for (int vi = g.vbeg; vi < g.vend; ++vi) {
                    const VisitQ34& v = vs[vi];
                    if (v.order <= 2) {
                        // fused order-1/2 adjoint: H C (1) and a complex
                        // FMA pair (4) per term; order 1 one FMA
                        f += v.order == 0 ? 0.0 : v.order == 1 ? 2.0 : 5.0 * (g.L + 1);
                        continue;
                    }
                    long cnt = 0;
                    for (int ma = -v.la; ma <= v.la; ++ma) {
                        if (v.order == 3) {
                            for (int k = 0; k <= g.L; ++k)
                                if (std::abs(k - ma) <= v.lb) ++cnt;
                        } else {
                            for (int mb = -v.lb; mb <= v.lb; ++mb)
                                for (int k = 0; k <= g.L; ++k)
                                    if (std::abs(k - ma - mb) <= v.lc) ++cnt;
                        }
                    }
                    // per term: complex product 6, then 4 (adjoint: two FMAs
                    // into G) or 5 (features: Re of the product with the
                    // output factor, one FMA); order 4 adds the a x b
                    // product (6) per (m_a, m_b), the adjoint H x a (2) per m_a.
                    // Visits of a species-paired group add their products
                    // into a visit accumulator instead and fold it into each
                    // of their rows, 4 (L + 1) per row
                    const double na_ = 2.0 * v.la + 1.0, nb_ = 2.0 * v.lb + 1.0;
                    const bool paired = adj && g.S == kPairedRows;
                    f += (adj ? 10.0 : 11.0) * cnt + (v.order == 4 ? 6.0 * na_ * nb_ : 0.0) +
                         (adj && !paired ? 2.0 * na_ : 0.0) + (paired ? (v.row == 2 ? 8.0 : 4.0) * (g.L + 1) : 0.0);
                }
            // adjoint: the energy of linear families from the G rows (Euler),
            // one row dot product per linear class present in the group
            if (adj)
                for (const GroupQ34& g : gs) {
                    bool seen[2][4] = {{false, false, false, false}, {false, false, false, false}};
                    for (int vi = g.vbeg; vi < g.vend; ++vi) {
                        const VisitQ34& v = vs[vi];
                        if (v.pslot >= 0 || v.order < 1 || v.order > 4) continue;
                        if (v.row != 1) seen[0][v.order - 1] = true;
                        if (g.S == kPairedRows && v.row != 0) seen[1][v.order - 1] = true;
                    }
                    for (int r = 0; r < 2; ++r)
                        for (bool b : seen[r]) f += b ? 5.0 * (g.L + 1) : 0.0;
                }
            return f;
        };
        out7[2] += na * terms(tt.groups_feat, tt.visits_feat, false);
        out7[3] += na * terms(tt.groups_adj, tt.visits_adj, true);
        // orders 5/6: per term the product of the other factors (6 per
        // complex product after the first), then 4 (adjoint: one complex
        // FMA pair into G) or 5 (features); per visit the G update and the
        // energy dot product, 10 (L + 1)
        auto hi_terms = [&](const std::vector<VisitHi>& vs, bool adj) {
            double f = 0.0;
            for (const VisitHi& v : vs)
                f += (v.tend - v.tbeg) * (6.0 * (v.order - 2) + (adj ? 4.0 : 5.0)) + (adj ? 10.0 * (v.L + 1) : 0.0);
            return f;
        };
        out7[2] += na * hi_terms(tt.hi_feat, false);
        out7[3] += na * hi_terms(tt.hi_adj, true);
        const double np = m.n_p, w = static_cast<double>(m.prange[t][1] - m.prange[t][0]) * m.n_fn;
        out7[4] += static_cast<double>(c->owners_per_type[t]) * (np > 0 ? 2.0 * w * w + 6.0 * w + 12.0 * m.c3_w.size() / np : 0.0);
    }
    for (int k = 0; k < 6; ++k) out7[6] += out7[k];
    return MLIP_OK;
}

// Algorithmic FLOPs of the current structure by the reference-formulation
// rule (counting weights: add or mul 1, FMA 2, complex x complex 6,
// complex x real 2; transcendental or division 1):
//   per half pair   40 + 12 n_fn + 38 n_lmh          geometry, radial, Y_lm
//                   + 8 n_fn n_lmh                    a_nlm accumulation
//                   + 44 n_fn n_lmh                   pair forces
//   per atom        n_fn sum_{active itc} T (6(q-1) + 3)     invariants
//                   n_fn sum_{active itc} T q (6(q-1) + 4)   adjoint G
//                   2 sum_{active poly terms} |p|            E and H
// T = CG terms of the itc's irrep, q its order.  Counted by the kernel that
// carries the work: a_nlm (geometry + accumulation), q12 (orders <= 2),
// q34_feat / q34_adj (orders >= 3: polynomial features / the rest with G),
// poly, force; out7[6] = total.
int mlip_flops_reference(const mlip_ctx* c, double* out7) {
    if (!c || !out7) return MLIP_ERR_ARG;
    const Model& m = c->model;
    for (int k = 0; k < 7; ++k) out7[k] = 0.0;
    const double half = 0.5 * static_cast<double>(c->npairs);
    const double nl = m.n_lmh, nf = m.n_fn;
    if (m.pair_model) {
        // radial-only style: geometry and radial functions, two feature
        // accumulations and the pair force per half pair, 2 sum|p| per atom
        double na = 0.0;
        for (int64_t k : c->owners_per_type) na += static_cast<double>(k);
        out7[0] = half * (40.0 + 12.0 * nf + 2.0 * nf);
        out7[5] = half * 16.0 * nf;
        out7[4] = na * 2.0 * (m.n_des + 2.0 * m.comb2.size() + 3.0 * m.comb3.size());
        for (int k = 0; k < 6; ++k) out7[6] += out7[k];
        return MLIP_OK;
    }
    out7[0] = half * (40.0 + 12.0 * nf + 38.0 * nl + 8.0 * nf * nl);
    out7[5] = half * 44.0 * nf * nl;
    std::vector<char> in_poly(m.n_gtinv, 0);
    if (m.maxp > 1)
        for (int f : m.poly_features) in_poly[f % m.n_gtinv] = 1;
    for (int t = 0; t < m.n_type && t < static_cast<int>(c->owners_per_type.size()); ++t) {
        const double na = static_cast<double>(c->owners_per_type[t]);
        if (na == 0.0) continue;
        for (int i = 0; i < m.n_gtinv; ++i) {
            const Itc& it = m.itcs[i];
            if (!((it.type_mask >> t) & 1u)) continue;
            const double T = static_cast<double>(m.irreps[it.irrep].m.size()), q = it.order;
            const double inv = nf * T * (6.0 * (q - 1.0) + 3.0), adj = nf * T * q * (6.0 * (q - 1.0) + 4.0);
            if (it.order <= 2)
                out7[1] += na * (inv + adj);
            else if (in_poly[i])
                out7[2] += na * inv, out7[3] += na * adj;
            else
                out7[3] += na * (inv + adj);
        }
        auto active = [&](int f) { return ((m.itcs[f % m.n_gtinv].type_mask >> t) & 1u) != 0; };
        double terms = 0.0;
        for (int f = 0; f < m.n_des; ++f) terms += active(f) ? 1.0 : 0.0;
        for (const auto& p2 : m.comb2) terms += (active(p2[0]) && active(p2[1])) ? 2.0 : 0.0;
        for (const auto& p3 : m.comb3) terms += (active(p3[0]) && active(p3[1]) && active(p3[2])) ? 3.0 : 0.0;
        out7[4] += na * 2.0 * terms;
    }
    for (int k = 0; k < 6; ++k) out7[6] += out7[k];
    return MLIP_OK;
}

// Plain device copy of nbytes (calibrates HBM byte counters against a
// kernel whose traffic is known exactly).
int mlip_calibration_copy(mlip_ctx* c, int64_t nbytes, int reps) {
    if (!c || nbytes <= 0) return MLIP_ERR_ARG;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, nbytes) != hipSuccess || hipMalloc(&b, nbytes) != hipSuccess) {
        if (a) (void)hipFree(a);
        c->err = "calibration buffers";
        return MLIP_ERR_DEVICE;
    }
    (void)hipMemsetAsync(a, 1, nbytes, c->stream);
    const bool ok = launch_copy(b, a, nbytes, reps, c->stream) == hipSuccess &&
                    hipStreamSynchronize(c->stream) == hipSuccess;
    (void)hipFree(a);
    (void)hipFree(b);
    return ok ? MLIP_OK : MLIP_ERR_DEVICE;
}

int64_t mlip_debug_fetch(mlip_ctx* c, int what, double* out, int64_t cap) {
    if (!c || !out) return -1;
    const Model& m = c->model;
    const size_t nslot = static_cast<size_t>(c->ds.nblocks) * kBlock;
    const double* src = nullptr;
    size_t n = 0;
    if (what == 0) {
        src = c->A.p;
        n = nslot * m.n_fn * m.n_type * m.n_lmh * 2;
    } else if (what == 1) {
        src = c->D.p;
        n = nslot * m.n_p;
    } else if (what == 2) {
        src = c->e_slot.p;
        n = nslot;
    } else if (what == 3) {
        src = c->G.p;
        n = nslot * m.n_fn * m.n_type * m.n_lmh * 2;
    } else if (what == 4) {
        // the quadratic-form GEMM's output: D W, or lin + D W (= H) where its
        // epilogue forms H (no cubic terms)
        src = c->D.p ? (c->dm.poly_fused ? c->H.p : c->Hq.p) : nullptr;
        n = nslot * m.n_p;
    } else if (what == 5) {
        src = c->d_wmat.p;
        n = c->wofs[1] + static_cast<size_t>(c->wdim[1]) * c->wdim[1];
    } else if (what == 6) {
        // layout of the per-species products: row pitch, then per species
        // (offset of its quadratic form, its dimension, its slot blocks)
        const double v[8] = {static_cast<double>(c->dm.n_pr), static_cast<double>(c->wofs[0]),
                             static_cast<double>(c->wdim[0]), static_cast<double>(c->nblk_type[0]),
                             static_cast<double>(c->wofs[1]), static_cast<double>(c->wdim[1]),
                             static_cast<double>(c->nblk_type[1]), static_cast<double>(c->dm.poly_fused)};
        if (cap < 8) return -1;
        std::copy(v, v + 8, out);
        return 8;
    } else if (what == 7) {
        src = c->d_linp.p;  // linear weights per species column ([2][row pitch])
        n = static_cast<size_t>(2) * c->dm.n_pr;
    }
    if (!src || static_cast<int64_t>(n) > cap) return -1;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
    if (hipMemcpy(out, src, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return static_cast<int64_t>(n);
}

int mlip_model_inspect(const char* potfile, int which, int* out, int64_t cap, char* errbuf, size_t errlen) {
    GtinvTable table;
    Model m;
    std::string e;
    if (!potfile || !out || !load_builtin_gtinv(table, e) || !load_model(potfile, table, m, e, which != 4)) {
        set_err(errbuf, errlen, e.empty() ? "bad arguments" : e);
        return -1;
    }
    std::vector<int> v;
    if (which == 4) {
        v = {m.n_coeff};
    } else if (which == 0) {
        v = {m.n_type, m.n_fn, m.n_gtinv, m.n_des, m.n_coeff, m.maxl, m.lmax3, m.n_p,
             static_cast<int>(m.irreps.size()), m.gtinv_order, m.model_type, m.maxp};
    } else if (which == 1) {
        for (const Itc& it : m.itcs) {
            v.push_back(it.irrep);
            v.push_back(it.order);
            for (int k = 0; k < 6; ++k) v.push_back(it.tc[k]);
            v.push_back(static_cast<int>(it.type_mask));
        }
    } else if (which == 2) {
        for (const auto& c2 : m.comb2) v.insert(v.end(), c2.begin(), c2.end());
    } else if (which == 3) {
        for (const auto& c3 : m.comb3) v.insert(v.end(), c3.begin(), c3.end());
    } else {
        set_err(errbuf, errlen, "unknown table");
        return -1;
    }
    if (static_cast<int64_t>(v.size()) > cap) {
        set_err(errbuf, errlen, "buffer too small");
        return -2;
    }
    std::copy(v.begin(), v.end(), out);
    return static_cast<int>(v.size());
}

struct mlip_nlist {
    NeighborList nl;
};

int mlip_nlist_build(const double* cell9, int natoms, const double* x, double rc, mlip_nlist** out, char* errbuf,
                     size_t errlen) {
    if (!cell9 || !x || !out || natoms < 0) {
        set_err(errbuf, errlen, "bad arguments");
        return MLIP_ERR_ARG;
    }
    auto* h = new mlip_nlist();
    std::string e;
    if (!build_full_list(cell9, natoms, x, rc, h->nl, e)) {
        set_err(errbuf, errlen, e);
        delete h;
        return MLIP_ERR_ARG;
    }
    *out = h;
    return MLIP_OK;
}

int64_t mlip_nlist_npairs(const mlip_nlist* h) { return h ? static_cast<int64_t>(h->nl.idx.size()) : -1; }

int mlip_nlist_copy(const mlip_nlist* h, int64_t* off, int* idx, double* shift) {
    if (!h) return MLIP_ERR_ARG;
    if (off) std::copy(h->nl.off.begin(), h->nl.off.end(), off);
    if (idx) std::copy(h->nl.idx.begin(), h->nl.idx.end(), idx);
    if (shift) std::copy(h->nl.shift.begin(), h->nl.shift.end(), shift);
    return MLIP_OK;
}

void mlip_nlist_free(mlip_nlist* h) { delete h; }

int mlip_model_info(const mlip_ctx* c, int* info) {
    if (!c || !info) return MLIP_ERR_ARG;
    const Model& m = c->model;
    int v[12] = {m.n_type, m.n_fn, m.n_gtinv, m.n_des, m.n_coeff, m.maxl, m.lmax3, m.n_p,
                 static_cast<int>(m.irreps.size()), m.gtinv_order, m.model_type, m.maxp};
    std::memcpy(info, v, sizeof v);
    return MLIP_OK;
}

}  // extern "C"
