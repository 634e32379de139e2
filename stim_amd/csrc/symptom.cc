// Host side of the sparse-symptom engine: groups noise sites into classes with one probability and
// outcome table, and lays out one symptom column per basis fault. See symptom.h.
#include "symptom.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>

namespace sa {

namespace {

uint32_t basis_of(uint32_t code) {
    switch (code) {
        case OP_NOISE1: return 2;   // X, Z
        case OP_NOISE2: return 4;   // X1, Z1, X2, Z2
        case OP_MFLIP: return 1;    // record flip
        case OP_HERALD: return 3;   // X, Z, herald
        default: return 0;
    }
}

}  // namespace

SymPlan plan_symptoms(const Program &prog) {
    SymPlan plan;
    const size_t nops = prog.ops.size();
    plan.op_class.assign(nops, 0xFFFFFFFFu);
    plan.op_site0.assign(nops, 0);
    plan.op_link.assign(nops, 0);

    struct Member {
        uint32_t op;
    };
    // Classes keyed by (opcode, p bits, outcome table words).
    std::map<std::vector<uint32_t>, uint32_t> key_to_class;
    std::vector<std::vector<uint32_t>> class_ops;   // member ops (site classes) or chain links
    std::vector<bool> class_is_chain;
    std::vector<std::vector<std::pair<double, uint32_t>>> chain_outs;

    auto add_chain = [&](const std::vector<uint32_t> &links) {
        if (links.empty()) return;
        if (links.size() > 22) {
            plan.supported = false;
            plan.why_unsupported = "correlated-error chain longer than 22 links";
            return;
        }
        // Link k fires with probability p_k * prod_{j<k} (1 - p_j) (frame_simulator.inl:750-776).
        std::vector<std::pair<double, uint32_t>> outs;
        double none = 1.0;
        for (size_t k = 0; k < links.size(); k++) {
            const double pk = prog.ops[links[k]].p;
            outs.push_back({none * pk, 1u << k});
            none *= 1.0 - pk;
        }
        uint32_t c = (uint32_t)class_ops.size();
        class_ops.push_back(links);
        class_is_chain.push_back(true);
        chain_outs.push_back(outs);
        for (size_t k = 0; k < links.size(); k++) {
            plan.op_class[links[k]] = c;
            plan.op_link[links[k]] = (uint32_t)k;
        }
    };

    std::vector<uint32_t> chain;
    for (uint32_t i = 0; i < nops; i++) {
        const DevOp &op = prog.ops[i];
        const uint32_t code = op.code & 0xFF;
        if (code == OP_CHAINCLR) {
            add_chain(chain);
            chain.clear();
            continue;
        }
        if (code == OP_CORR) {
            chain.push_back(i);
            continue;
        }
        const uint32_t b = basis_of(code);
        if (!b || !(op.p > 0.0f)) continue;
        std::vector<uint32_t> key{code};
        uint32_t pb;
        std::memcpy(&pb, &op.p, 4);
        key.push_back(pb);
        const uint32_t nout = op.nout & 0xFFFF;
        for (uint32_t k = 0; k < 2 * nout; k++) key.push_back(prog.otab[op.woff + k]);
        auto it = key_to_class.find(key);
        uint32_t c;
        if (it == key_to_class.end()) {
            c = (uint32_t)class_ops.size();
            key_to_class[key] = c;
            class_ops.push_back({});
            class_is_chain.push_back(false);
            chain_outs.push_back({});
        } else {
            c = it->second;
        }
        class_ops[c].push_back(i);
        plan.op_class[i] = c;
    }
    add_chain(chain);

    // Lay out classes, sites and columns, in order of firing probability so that classes with equal
    // probability are adjacent (the sampler draws them as one process).
    auto class_p = [&](uint32_t c) -> float {
        if (!class_is_chain[c]) return prog.ops[class_ops[c][0]].p;
        double none = 1.0;
        for (uint32_t l : class_ops[c]) none *= 1.0 - (double)prog.ops[l].p;
        return (float)(1.0 - none);
    };
    std::vector<uint32_t> order(class_ops.size());
    for (uint32_t c = 0; c < order.size(); c++) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return class_p(a) < class_p(b); });
    std::vector<uint32_t> new_index(class_ops.size());
    for (uint32_t k = 0; k < order.size(); k++) new_index[order[k]] = k;
    for (uint32_t &c : plan.op_class) {
        if (c != 0xFFFFFFFFu) c = new_index[c];
    }
    uint32_t col = 0;
    uint64_t sites = 0;
    for (uint32_t c : order) {
        SymClass sc{};
        sc.col0 = col;
        sc.site0 = (uint32_t)sites;
        sc.woff = (uint32_t)plan.otab.size();
        if (class_is_chain[c]) {
            const auto &links = class_ops[c];
            double none = 1.0;
            for (uint32_t l : links) none *= 1.0 - (double)prog.ops[l].p;
            const double p = 1.0 - none;
            sc.p = (float)p;
            sc.inv_log = (p > 0 && p < 1) ? (float)(1.0 / std::log2(1.0 - p)) : 0.0f;
            sc.nsites = 1;
            sc.basis = (uint32_t)links.size();
            double tot = 0, acc = 0;
            for (auto &o : chain_outs[c]) tot += o.first;
            for (size_t k = 0; k < chain_outs[c].size(); k++) {
                acc += chain_outs[c][k].first;
                double thr = tot > 0 ? acc / tot * 4294967296.0 - 1.0 : 4294967295.0;
                plan.otab.push_back(k + 1 == chain_outs[c].size() || thr >= 4294967295.0 ? 0xFFFFFFFFu
                                                                                          : (uint32_t)std::max(0.0, thr));
                plan.otab.push_back(chain_outs[c][k].second);
            }
            sc.nout = (uint32_t)chain_outs[c].size();
            for (size_t k = 0; k < links.size(); k++) {
                plan.columns.push_back({links[k], 0, 1u});
                plan.op_site0[links[k]] = 0;
            }
            col += sc.basis;
            sites += 1;
        } else {
            const DevOp &first = prog.ops[class_ops[c][0]];
            const uint32_t code = first.code & 0xFF;
            sc.p = first.p;
            sc.inv_log = first.inv_log;
            sc.basis = basis_of(code);
            sc.nout = first.nout & 0xFFFF;
            for (uint32_t k = 0; k < 2 * sc.nout; k++) plan.otab.push_back(prog.otab[first.woff + k]);
            // Equal outcome weights (DEPOLARIZE1/2, HERALDED_ERASE, single-outcome channels) let the
            // kernel index the outcome directly instead of scanning thresholds.
            {
                bool uniform = true;
                for (uint32_t k = 0; k + 1 < sc.nout; k++) {
                    const double want = (double)(k + 1) / sc.nout * 4294967296.0 - 1.0;
                    if (std::fabs((double)prog.otab[first.woff + 2 * k] - want) > 2.0) uniform = false;
                }
                if (uniform) sc.nout |= SYM_UNIFORM;
                // Uniform over the codes 1..nout: the kernel computes the code without the table.
                bool seq = uniform;
                std::vector<bool> seen(sc.nout & 0xFFFF, false);
                for (uint32_t k = 0; seq && k < (sc.nout & 0xFFFF); k++) {
                    const uint32_t code = prog.otab[first.woff + 2 * k + 1];
                    if (code < 1 || code > (sc.nout & 0xFFFF) || seen[code - 1]) seq = false;
                    else seen[code - 1] = true;
                }
                if (seq) sc.nout |= SYM_SEQ;
            }
            uint32_t n = 0;
            for (uint32_t oi : class_ops[c]) {
                const DevOp &op = prog.ops[oi];
                plan.op_site0[oi] = n;
                for (uint32_t s = 0; s < op.n; s++) {
                    for (uint32_t j = 0; j < sc.basis; j++) plan.columns.push_back({oi, s, 1u << j});
                }
                n += op.n;
            }
            sc.nsites = n;
            col += n * sc.basis;
            sites += n;
        }
        plan.classes.push_back(sc);
    }
    plan.total_sites = sites;
    // Compact record slots: each class starts at a multiple of its stride, sites are stride apart.
    uint64_t slot = 0;
    for (SymClass &sc : plan.classes) {
        sc.rstride = sc.basis <= 4 ? 4 : sc.basis;
        const uint32_t align = sc.rstride == 4 ? 4 : 1;
        slot = (slot + align - 1) / align * align;
        sc.rec0 = (uint32_t)slot;
        slot += (uint64_t)sc.nsites * sc.rstride;
    }
    plan.total_slots = slot;
    // Groups: runs of up to 4 adjacent classes with equal probability and at most 4 basis columns,
    // sampled as one Bernoulli process over their sites; every other class is a group of its own.
    for (uint32_t c = 0; c < plan.classes.size(); c++) {
        const SymClass &sc = plan.classes[c];
        SymGroup *g = plan.groups.empty() ? nullptr : &plan.groups.back();
        const bool joinable = sc.basis <= 4;
        if (g && joinable && g->nmem < 4 && g->rstride == 4 && g->p == sc.p && g->inv_log == sc.inv_log) {
            g->bound[g->nmem] = sc.site0 - g->site0;
            g->nout[g->nmem] = sc.nout;
            g->woff[g->nmem] = sc.woff;
            g->nmem++;
            g->nsites += sc.nsites;
            continue;
        }
        SymGroup ng{};
        ng.p = sc.p;
        ng.inv_log = sc.inv_log;
        ng.site0 = sc.site0;
        ng.nsites = sc.nsites;
        ng.rec0 = sc.rec0;
        ng.rstride = joinable ? 4 : sc.rstride;
        ng.nmem = 1;
        ng.c0 = c;
        for (int k = 0; k < 4; k++) ng.bound[k] = 0xFFFFFFFFu;
        ng.bound[0] = 0;
        ng.nout[0] = sc.nout;
        ng.woff[0] = sc.woff;
        ng.flags = 0;
        plan.groups.push_back(ng);
    }
    if (plan.columns.size() >= (1ull << 31) || plan.total_slots >= (1ull << 31)) {
        plan.supported = false;
        plan.why_unsupported = "too many noise sites for the symptom table";
    }
    return plan;
}

}  // namespace sa
