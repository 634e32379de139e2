// Lowers a REPEAT-unrolled circuit into the device program of program.h.
//
// Gate -> rule mapping follows FrameSimulator::do_gate (src/stim/simulators/frame_simulator.inl:914-1113).
// MPP / MXX / MYY / MZZ become Pauli-product measurements whose recorded flip is the frame's
// anticommutation parity with the product; this is what the reference's decomposition into basis
// changes + CX fan-in + M + uncompute (src/stim/circuit/gate_decomposition.cc:88-170, 245-274;
// frame_simulator.inl:797-840) does to a Pauli frame. SPP/SPP_DAG multiply the frame by the product
// when it anticommutes (decompose_spp_or_spp_dag_operation, gate_decomposition.cc:172-243).
// PAULI_CHANNEL_1/2 outcomes follow the disjoint order of perform_pauli_errors_via_correlated_errors
// (src/stim/simulators/tableau_simulator.h:291-324); DEPOLARIZE1/2 codes follow do_DEPOLARIZE1/2
// (frame_simulator.inl:632-663).
//
// Layout rules the kernel relies on:
//   * no two items of one frame op touch the same qubit (instructions are cut into segments);
//   * record slots written by one op are contiguous (ops are cut at the ring wrap);
//   * a frame op's words fit one LDS staging buffer (ops are capped at STAGE_WORDS words) and an
//     epoch (ops between barriers) fits one staging buffer;
//   * CSR headers hold offsets relative to the op's first word.
#include "compiler.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>

namespace sa {

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;

// Read / write / atomic access epochs per resource, used to place barriers between ops.
struct Tracker {
    std::vector<uint32_t> R, W, A;
    void resize(size_t n) {
        R.assign(n, 0);
        W.assign(n, 0);
        A.assign(n, 0);
    }
};

enum Kind : char { KR = 'r', KW = 'w', KA = 'a' };

struct Builder {
    Program P;
    std::vector<uint32_t> op_tlen;  // staged words per op (0 for noise / wordless ops)
    std::vector<uint32_t> qmap;
    uint64_t mcount = 0;
    uint32_t noise_count = 0;
    uint32_t epoch = 1;
    Tracker tq, tr, tm;  // qubits, record slots, misc (staging rows 0..255, obs 256, chain 257)
    std::vector<std::pair<uint32_t, char>> aq, ar, am;
    std::vector<uint32_t> seg_mark;
    uint32_t seg_id = 1;
    std::vector<std::vector<uint32_t>> det_terms;
    uint32_t det_base = 0;
    std::vector<std::vector<uint32_t>> obs_items;

    static constexpr uint32_t MISC_OBS = STAGE_ROWS;
    static constexpr uint32_t MISC_CHAIN = STAGE_ROWS + 1;

    uint32_t q(const Target &t) const {
        return qmap[t.value];
    }
    uint32_t slot_of(uint64_t m) const {
        return (uint32_t)(m % P.ring);
    }
    uint32_t lookback_slot(uint32_t k) const {
        if (k == 0 || k > mcount) {
            throw std::out_of_range("Referred to a measurement record before the beginning of time.");
        }
        return slot_of(mcount - k);
    }

    // ------------------------------------------------------------------ hazards
    void touch_q(uint32_t r, char k) { aq.push_back({r, k}); }
    void touch_r(uint32_t r, char k) { ar.push_back({r, k}); }
    void touch_m(uint32_t r, char k) { am.push_back({r, k}); }

    bool conflict(const Tracker &t, const std::vector<std::pair<uint32_t, char>> &acc) const {
        for (const auto &a : acc) {
            uint32_t r = a.first;
            bool w = t.W[r] == epoch, rd = t.R[r] == epoch, at = t.A[r] == epoch;
            if (a.second == KW && (w || rd || at)) return true;
            if (a.second == KR && (w || at)) return true;
            if (a.second == KA && (w || rd)) return true;
        }
        return false;
    }
    void mark(Tracker &t, const std::vector<std::pair<uint32_t, char>> &acc) {
        for (const auto &a : acc) {
            if (a.second == KW) t.W[a.first] = epoch;
            else if (a.second == KR) t.R[a.first] = epoch;
            else t.A[a.first] = epoch;
        }
    }
    void barrier_before_next() {
        if (!P.ops.empty() && !(P.ops.back().code & OPF_SYNC)) {
            P.ops.back().code |= OPF_SYNC;
            P.num_syncs++;
        }
        epoch++;
    }
    // Appends an op. Frame ops must have appended their words to P.tgt since make().
    void push(DevOp op, bool frame_words = true) {
        if (!P.ops.empty() && (conflict(tq, aq) || conflict(tr, ar) || conflict(tm, am))) {
            barrier_before_next();
        }
        mark(tq, aq);
        mark(tr, ar);
        mark(tm, am);
        aq.clear();
        ar.clear();
        am.clear();
        P.ops.push_back(op);
        op_tlen.push_back(frame_words ? (uint32_t)(P.tgt.size() - op.toff) : 0u);
    }
    DevOp make(uint32_t code, uint32_t n) {
        DevOp op{};
        op.code = code;
        op.n = n;
        op.toff = (uint32_t)P.tgt.size();
        return op;
    }

    // ------------------------------------------------------------------ segments
    bool seg_conflict(std::initializer_list<uint32_t> qs) {
        for (uint32_t x : qs) {
            if (seg_mark[x] == seg_id) return true;
        }
        return false;
    }
    void seg_add(std::initializer_list<uint32_t> qs) {
        for (uint32_t x : qs) seg_mark[x] = seg_id;
    }
    void seg_reset() { seg_id++; }

    // ------------------------------------------------------------------ emitters
    void emit_1q(uint32_t code, const std::vector<uint32_t> &qs, uint32_t flags = 0) {
        size_t i = 0;
        while (i < qs.size()) {
            seg_reset();
            DevOp op = make(code | flags, 0);
            while (i < qs.size() && op.n < OP_WORDS_MAX && !seg_conflict({qs[i]})) {
                seg_add({qs[i]});
                P.tgt.push_back(qs[i]);
                touch_q(qs[i], KW);
                op.n++;
                i++;
            }
            push(op);
        }
    }

    void emit_2q(uint32_t code, const std::vector<std::pair<uint32_t, uint32_t>> &ps) {
        size_t i = 0;
        while (i < ps.size()) {
            seg_reset();
            DevOp op = make(code, 0);
            while (i < ps.size() && 2 * op.n + 2 <= OP_WORDS_MAX && !seg_conflict({ps[i].first, ps[i].second})) {
                seg_add({ps[i].first, ps[i].second});
                P.tgt.push_back(ps[i].first);
                P.tgt.push_back(ps[i].second);
                touch_q(ps[i].first, KW);
                touch_q(ps[i].second, KW);
                op.n++;
                i++;
            }
            push(op);
        }
    }

    // Feedback items (slot, qubit, mask).
    void emit_feedback(const std::vector<std::array<uint32_t, 3>> &items) {
        size_t i = 0;
        while (i < items.size()) {
            seg_reset();
            DevOp op = make(OP_FEEDBACK, 0);
            while (i < items.size() && 3 * op.n + 3 <= OP_WORDS_MAX && !seg_conflict({items[i][1]})) {
                seg_add({items[i][1]});
                for (int k = 0; k < 3; k++) P.tgt.push_back(items[i][k]);
                touch_q(items[i][1], KW);
                touch_r(items[i][0], KR);
                op.n++;
                i++;
            }
            push(op);
        }
    }

    // Noise op builder; its site words must already be in P.ntgt starting at ntoff. Returns op index.
    uint32_t emit_noise(uint32_t code, uint32_t n_sites, uint32_t ntoff, double p,
                        const std::vector<std::pair<double, uint32_t>> &outs, uint32_t flat_k, uint32_t first_site,
                        uint32_t flags = 0) {
        DevOp op{};
        op.code = code | flags;
        op.n = n_sites;
        op.toff = ntoff;
        op.p = (float)p;
        op.inv_log = (p > 0 && p < 1) ? (float)(1.0 / std::log2(1.0 - p)) : 0.0f;
        // Outcome table: cumulative thresholds over uint32 in proportion to outcome weights.
        op.woff = (uint32_t)P.otab.size();
        double tot = 0;
        for (auto &o : outs) tot += o.first;
        double acc = 0;
        uint32_t nout = 0;
        for (size_t i = 0; i < outs.size(); i++) {
            if (outs[i].first <= 0 && outs.size() > 1) continue;
            acc += outs[i].first;
            double thr = tot > 0 ? acc / tot * 4294967296.0 - 1.0 : 4294967295.0;
            uint32_t t = thr >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)std::max(0.0, thr);
            P.otab.push_back(t);
            P.otab.push_back(outs[i].second);
            nout++;
        }
        if (nout == 0) {
            P.otab.push_back(0xFFFFFFFFu);
            P.otab.push_back(outs.empty() ? 0 : outs[0].second);
            nout = 1;
        }
        P.otab[P.otab.size() - 2] = 0xFFFFFFFFu;
        uint32_t chunk_log2 = 6;
        if (p > 0 && p < 1) {
            double want = std::ceil(1.0 / p);
            while (chunk_log2 < 20 && (double)(1u << chunk_log2) < want) chunk_log2++;
        }
        if (p > 0.3) op.code |= OPF_DENSE;
        op.nout = nout | (chunk_log2 << 16);
        noise_count++;
        if (flat_k != NONE) {
            P.sites[flat_k].push_back({(uint32_t)P.ops.size(), first_site, n_sites});
        }
        push(op, false);
        return (uint32_t)P.ops.size() - 1;
    }

    // ------------------------------------------------------------------ CSR ops
    // Emits ops of CSR items (header of n+1 offsets relative to the op's first word, then terms),
    // cut so each op's words fit one staging buffer. item_words[i] lists item i's words;
    // touch(i) records its accesses. 'prologue' runs before each op (e.g. record base).
    template <class TouchF, class OpF>
    void emit_csr(uint32_t code, const std::vector<std::vector<uint32_t>> &items, TouchF touch, OpF configure,
                  const std::function<bool(size_t, size_t)> &can_join = nullptr) {
        size_t i = 0;
        while (i < items.size()) {
            seg_reset();
            size_t e = i;
            size_t words = 1;
            while (e < items.size()) {
                size_t add = 1 + items[e].size();
                if (words + add > OP_WORDS_MAX) break;
                if (can_join && !can_join(i, e)) break;
                words += add;
                e++;
            }
            if (e == i) {
                throw std::invalid_argument("a detector, observable or Pauli product has too many terms for the GPU "
                                            "sampler (limit " + std::to_string(OP_WORDS_MAX - 2) + ")");
            }
            DevOp op = make(code, (uint32_t)(e - i));
            configure(op, i);
            const size_t cnt = e - i;
            const size_t hdr = P.tgt.size();
            P.tgt.resize(hdr + cnt + 1);
            for (size_t d = 0; d < cnt; d++) {
                P.tgt[hdr + d] = (uint32_t)(P.tgt.size() - op.toff);
                for (uint32_t v : items[i + d]) P.tgt.push_back(v);
                touch(i + d);
            }
            P.tgt[hdr + cnt] = (uint32_t)(P.tgt.size() - op.toff);
            push(op);
            i = e;
        }
    }

    // ------------------------------------------------------------------ detectors / observables
    void flush_pending_dets() {
        size_t i = 0;
        while (i < det_terms.size()) {
            uint32_t room = STAGE_ROWS - (det_base % STAGE_ROWS);
            size_t cnt = std::min<size_t>(room, det_terms.size() - i);
            std::vector<std::vector<uint32_t>> chunk(det_terms.begin() + (long)i, det_terms.begin() + (long)(i + cnt));
            size_t done = 0;
            emit_csr(
                OP_DET, chunk,
                [&](size_t d) {
                    for (uint32_t s : chunk[d]) touch_r(s, KR);
                    touch_m((det_base + (uint32_t)d) % STAGE_ROWS, KW);
                },
                [&](DevOp &op, size_t first) {
                    op.base = det_base + (uint32_t)first;
                    done = first;
                });
            (void)done;
            det_base += (uint32_t)cnt;
            i += cnt;
            if (det_base % STAGE_ROWS == 0) {
                emit_flush(det_base - STAGE_ROWS, STAGE_ROWS);
            }
        }
        det_terms.clear();
    }
    void emit_flush(uint32_t base, uint32_t n) {
        DevOp op = make(OP_FLUSH, n);
        op.base = base;
        for (uint32_t r = 0; r < n; r++) touch_m(r, KR);
        push(op);
    }
    void flush_pending_obs() {
        if (obs_items.empty()) return;
        emit_csr(
            OP_OBS, obs_items,
            [&](size_t i) {
                for (size_t k = 1; k < obs_items[i].size(); k++) {
                    uint32_t t = obs_items[i][k];
                    if (t & TERM_PAULI) touch_q(t & TERM_QMASK, KR);
                    else touch_r(t, KR);
                }
                touch_m(MISC_OBS, KA);
            },
            [&](DevOp &, size_t) {});
        obs_items.clear();
    }
    void flush_pending() {
        flush_pending_dets();
        flush_pending_obs();
    }

    // ------------------------------------------------------------------ measurements
    // Record-producing op over items, cut at ring wraps so slots stay contiguous, at qubit
    // repeats, and at the staging size.
    template <class F>
    void emit_records(uint32_t code, size_t n, F body, const std::function<std::vector<uint32_t>(size_t)> &qubits_of,
                      uint32_t flags, std::vector<std::pair<uint32_t, uint32_t>> *ranges) {
        size_t i = 0;
        while (i < n) {
            seg_reset();
            DevOp op = make(code | flags, 0);
            op.base = slot_of(mcount);
            size_t first = i;
            while (i < n && op.n < OP_WORDS_MAX) {
                if (slot_of(mcount) == 0 && i > first) break;  // ring wrap
                if (qubits_of) {
                    std::vector<uint32_t> qs = qubits_of(i);
                    bool c = false;
                    for (uint32_t x : qs) c |= seg_mark[x] == seg_id;
                    if (c) break;
                    for (uint32_t x : qs) seg_mark[x] = seg_id;
                }
                body(i);
                touch_r(slot_of(mcount), KW);
                mcount++;
                op.n++;
                i++;
            }
            push(op);
            if (ranges) ranges->push_back({(uint32_t)first, (uint32_t)(i - first)});
        }
    }

    // Record flips of noisy results, one MFLIP op per contiguous slot range.
    void emit_mflips(double p, uint64_t m_first, const std::vector<std::pair<uint32_t, uint32_t>> &ranges, uint32_t k) {
        for (const auto &rg : ranges) {
            uint64_t m0 = m_first + rg.first;
            for (uint32_t j = 0; j < rg.second; j++) touch_r(slot_of(m0 + j), KA);
            uint32_t opi = emit_noise(OP_MFLIP, rg.second, (uint32_t)P.ntgt.size(), p, {{1.0, 1u}}, k, rg.first);
            P.ops[opi].base = slot_of(m0);
        }
    }

    void measure_1q(const Instr &in, uint32_t k, uint32_t code) {
        flush_pending();
        std::vector<uint32_t> qs;
        for (const Target &t : in.targets) qs.push_back(q(t));
        uint64_t m_first = mcount;
        std::vector<std::pair<uint32_t, uint32_t>> ranges;
        emit_records(
            code, qs.size(),
            [&](size_t i) {
                P.tgt.push_back(qs[i]);
                touch_q(qs[i], KW);
            },
            [&](size_t i) { return std::vector<uint32_t>{qs[i]}; }, OPF_RANDOMIZE, &ranges);
        P.has_randomizing_ops = true;
        if (!in.args.empty()) emit_mflips(in.args[0], m_first, ranges, k);
    }

    void measure_products(const Instr &in, uint32_t k, const std::vector<std::vector<uint32_t>> &products) {
        flush_pending();
        uint64_t m_first = mcount;
        std::vector<std::pair<uint32_t, uint32_t>> ranges;
        // Products are cut where they overlap an earlier product of the same op (they are measured
        // one after another), at ring wraps, and at the staging size.
        std::vector<uint32_t> seg_of(products.size());
        {
            seg_reset();
            uint32_t cur = seg_id;
            for (size_t i = 0; i < products.size(); i++) {
                bool c = false;
                for (uint32_t t : products[i]) c |= seg_mark[t & TERM_QMASK] == seg_id;
                if (c) {
                    seg_reset();
                    cur = seg_id;
                }
                for (uint32_t t : products[i]) seg_mark[t & TERM_QMASK] = seg_id;
                seg_of[i] = cur;
            }
        }
        size_t placed = 0;
        emit_csr(
            OP_MPP | OPF_RANDOMIZE, products,
            [&](size_t i) {
                for (uint32_t t : products[i]) touch_q(t & TERM_QMASK, KW);
                touch_r(slot_of(mcount), KW);
                mcount++;
                placed = i + 1;
            },
            [&](DevOp &op, size_t first) {
                op.base = slot_of(mcount);
                if (!ranges.empty() || first) {
                }
                ranges.push_back({(uint32_t)first, 0});
            },
            [&](size_t first, size_t e) {
                if (e == first) return true;
                if (seg_of[e] != seg_of[first]) return false;
                return slot_of(m_first + e) != 0;  // ring wrap
            });
        // Fill range counts from consecutive starts.
        for (size_t r = 0; r < ranges.size(); r++) {
            size_t end = r + 1 < ranges.size() ? ranges[r + 1].first : products.size();
            ranges[r].second = (uint32_t)(end - ranges[r].first);
        }
        (void)placed;
        P.has_randomizing_ops = true;
        if (!in.args.empty()) emit_mflips(in.args[0], m_first, ranges, k);
    }

    // Combines '*'-joined Pauli terms into a product; throws on anti-Hermitian products.
    std::vector<uint32_t> combine_product(const Instr &in, size_t b, size_t e, std::vector<uint32_t> *bits) {
        std::vector<std::pair<uint32_t, uint8_t>> acc;
        int imag = 0;
        for (size_t i = b; i < e; i++) {
            const Target &t = in.targets[i];
            if (t.kind == TK::COMBINER) continue;
            if (t.kind == TK::PAULI) {
                uint32_t dq = q(t);
                bool found = false;
                for (auto &a : acc) {
                    if (a.first == dq) {
                        if (a.second && a.second != t.pauli) imag ^= 1;
                        a.second ^= t.pauli;
                        found = true;
                    }
                }
                if (!found) acc.push_back({dq, t.pauli});
            } else if ((t.kind == TK::REC || t.kind == TK::SWEEP) && bits) {
                if (t.kind == TK::REC) bits->push_back(lookback_slot(t.value));
            } else {
                throw std::invalid_argument("Found an unsupported target in a Pauli product.");
            }
        }
        if (imag) {
            throw std::invalid_argument("Acted on an anti-Hermitian operator (e.g. X0*Z0 instead of Y0).");
        }
        std::vector<uint32_t> terms;
        for (auto &a : acc) {
            if (!a.second) continue;
            terms.push_back(TERM_PAULI | a.first | ((a.second & 1) ? TERM_X : 0) | ((a.second & 2) ? TERM_Z : 0));
        }
        return terms;
    }

    // ------------------------------------------------------------------ main dispatch
    // Measurement outputs: the circuit's detectors and observables are dropped and every measurement
    // result becomes an output row of its own (a one-term detector right after the instruction).
    bool meas_outputs = false;

    void run(const std::vector<const Instr *> &flat) {
        P.stats = compute_stats(flat);
        if (meas_outputs) {
            // The ring keeps every result of one instruction until its output rows are read.
            uint64_t widest = 0;
            for (const Instr *in : flat) widest = std::max<uint64_t>(widest, num_results(*in));
            P.stats.max_lookback = std::max<uint64_t>(P.stats.max_lookback, widest);
            P.stats.num_detectors = P.stats.num_measurements;
            P.stats.num_observables = 0;
        }
        std::vector<uint8_t> used(P.stats.num_qubits + 1, 0);
        for (const Instr *in : flat) {
            if (in->gate == Gate::MPAD || in->gate == Gate::QUBIT_COORDS) continue;
            for (const Target &t : in->targets) {
                if (t.kind == TK::QUBIT || t.kind == TK::PAULI) used[t.value] = 1;
            }
        }
        qmap.assign(used.size(), NONE);
        uint32_t nq = 0;
        for (size_t i = 0; i < used.size(); i++) {
            if (used[i]) qmap[i] = nq++;
        }
        P.num_qubits = std::max<uint32_t>(nq, 1);
        P.ring = (uint32_t)std::max<uint64_t>(P.stats.max_lookback, 256);
        if (P.stats.num_measurements >= (1ull << 32) || P.stats.num_detectors >= (1ull << 31)) {
            throw std::invalid_argument("circuit has too many measurements or detectors for the GPU sampler");
        }
        P.num_detectors = (uint32_t)P.stats.num_detectors;
        P.num_observables = (uint32_t)P.stats.num_observables;
        P.num_measurements = P.stats.num_measurements;
        P.sites.assign(flat.size(), {});
        tq.resize(P.num_qubits);
        tr.resize(P.ring);
        tm.resize(STAGE_ROWS + 2);
        seg_mark.assign(P.num_qubits, 0);

        for (uint32_t k = 0; k < (uint32_t)flat.size(); k++) {
            const Gate g = flat[k]->gate;
            if (meas_outputs && (g == Gate::DETECTOR || g == Gate::OBSERVABLE_INCLUDE)) continue;
            const uint64_t before = mcount;
            step(*flat[k], k);
            if (meas_outputs) {
                for (uint64_t j = mcount - before; j > 0; j--) det_terms.push_back({lookback_slot((uint32_t)j)});
            }
        }
        flush_pending();
        if (det_base % STAGE_ROWS) {
            emit_flush(det_base - det_base % STAGE_ROWS, det_base % STAGE_ROWS);
        }
        if (P.num_observables) {
            DevOp op = make(OP_FLUSH_OBS, P.num_observables);
            touch_m(MISC_OBS, KR);
            push(op);
        }
        if (!P.ops.empty() && !(P.ops.back().code & OPF_SYNC)) {
            P.ops.back().code |= OPF_SYNC;
            P.num_syncs++;
        }
        P.num_noise_ops = noise_count;
        if (P.tgt.size() >= (1ull << 32) || P.ntgt.size() >= (1ull << 32)) {
            throw std::invalid_argument("circuit program too large");
        }
        build_epochs();
    }

    // Groups ops into barrier-free runs whose blob (header + op headers + words) fits one staging
    // buffer, then packs the blobs.
    void build_epochs() {
        uint32_t b = 0;
        uint64_t tbegin = 0, tend = 0;
        bool any = false;
        auto close = [&](uint32_t e) {
            DevEpoch ep{};
            ep.op_begin = b;
            ep.op_end = e;
            ep.tbegin = any ? (uint32_t)tbegin : 0;
            ep.tlen = any ? (uint32_t)(tend - tbegin) : 0;
            P.epochs.push_back(ep);
            b = e;
            any = false;
        };
        for (uint32_t i = 0; i < (uint32_t)P.ops.size(); i++) {
            const uint32_t len = op_tlen[i];
            const uint64_t s = P.ops[i].toff, e = s + len;
            const uint64_t words = (uint64_t)BLOB_HEADER * (i - b + 2) + (len ? e - (any ? tbegin : s) : (any ? tend - tbegin : 0));
            if (i > b && words > STAGE_WORDS) {
                // Would overflow the staging buffer: barrier before this op.
                if (!(P.ops[i - 1].code & OPF_SYNC)) {
                    P.ops[i - 1].code |= OPF_SYNC;
                    P.num_syncs++;
                }
                close(i);
            }
            if (len) {
                if (!any) {
                    tbegin = s;
                    any = true;
                }
                tend = e;
            }
            if (P.ops[i].code & OPF_SYNC) close(i + 1);
        }
        if (b < P.ops.size()) close((uint32_t)P.ops.size());
        // Pack blobs.
        std::vector<uint32_t> offs(P.epochs.size() + 1, 0);
        for (size_t e = 0; e < P.epochs.size(); e++) {
            const DevEpoch &ep = P.epochs[e];
            offs[e + 1] = offs[e] + BLOB_HEADER * (1 + ep.op_end - ep.op_begin) + ep.tlen;
            if (offs[e + 1] - offs[e] > STAGE_WORDS) throw std::logic_error("epoch blob exceeds the staging buffer");
        }
        P.blob.assign(offs.back(), 0);
        P.blob0_len = P.epochs.empty() ? 0 : offs[1];
        for (size_t e = 0; e < P.epochs.size(); e++) {
            const DevEpoch &ep = P.epochs[e];
            uint32_t *blob = P.blob.data() + offs[e];
            const uint32_t nops = ep.op_end - ep.op_begin;
            blob[0] = nops;
            blob[1] = offs[e + 1];
            blob[2] = e + 1 < P.epochs.size() ? offs[e + 2] - offs[e + 1] : 0;
            blob[3] = ep.op_begin;
            const uint32_t words0 = BLOB_HEADER * (1 + nops);
            for (uint32_t j = 0; j < nops; j++) {
                DevOp op = P.ops[ep.op_begin + j];
                if (op_tlen[ep.op_begin + j] || !is_noise(op.code & 0xFF)) {
                    op.toff = op_tlen[ep.op_begin + j] ? words0 + (op.toff - ep.tbegin) : 0;
                }
                std::memcpy(blob + BLOB_HEADER * (1 + j), &op, sizeof(DevOp));
            }
            for (uint32_t k = 0; k < ep.tlen; k++) blob[words0 + k] = P.tgt[ep.tbegin + k];
        }
    }

    static bool is_noise(uint32_t code) {
        return code == OP_NOISE1 || code == OP_NOISE2 || code == OP_MFLIP || code == OP_CORR || code == OP_HERALD;
    }

    void step(const Instr &in, uint32_t k) {
        const auto &T = in.targets;
        switch (in.gate) {
            case Gate::TICK: case Gate::QUBIT_COORDS: case Gate::SHIFT_COORDS:
            case Gate::I: case Gate::X: case Gate::Y: case Gate::Z: case Gate::II:
            case Gate::I_ERROR: case Gate::II_ERROR:
                return;
            case Gate::DETECTOR: {
                flush_pending_obs();
                std::vector<uint32_t> terms;
                for (const Target &t : T) terms.push_back(lookback_slot(t.value));
                det_terms.push_back(std::move(terms));
                return;
            }
            case Gate::OBSERVABLE_INCLUDE: {
                flush_pending_dets();
                std::vector<uint32_t> item{(uint32_t)in.args[0]};
                for (const Target &t : T) {
                    if (t.kind == TK::REC) {
                        item.push_back(lookback_slot(t.value));
                    } else {
                        item.push_back(TERM_PAULI | q(t) | ((t.pauli & 1) ? TERM_X : 0) | ((t.pauli & 2) ? TERM_Z : 0));
                    }
                }
                obs_items.push_back(std::move(item));
                return;
            }
            default:
                break;
        }
        flush_pending();
        auto qubits = [&]() {
            std::vector<uint32_t> v;
            for (const Target &t : T) v.push_back(q(t));
            return v;
        };
        switch (in.gate) {
            case Gate::H: case Gate::H_NXZ: case Gate::SQRT_Y: case Gate::SQRT_Y_DAG:
                emit_1q(OP_H, qubits());
                return;
            case Gate::S: case Gate::S_DAG: case Gate::H_XY: case Gate::H_NXY:
                emit_1q(OP_HXY, qubits());
                return;
            case Gate::SQRT_X: case Gate::SQRT_X_DAG: case Gate::H_YZ: case Gate::H_NYZ:
                emit_1q(OP_HYZ, qubits());
                return;
            case Gate::C_XYZ: case Gate::C_NXYZ: case Gate::C_XNYZ: case Gate::C_XYNZ:
                emit_1q(OP_CXYZ, qubits());
                return;
            case Gate::C_ZYX: case Gate::C_NZYX: case Gate::C_ZNYX: case Gate::C_ZYNX:
                emit_1q(OP_CZYX, qubits());
                return;
            case Gate::R:
                emit_1q(OP_R, qubits(), OPF_RANDOMIZE);
                P.has_randomizing_ops = true;
                return;
            case Gate::RX:
                emit_1q(OP_RX, qubits(), OPF_RANDOMIZE);
                P.has_randomizing_ops = true;
                return;
            case Gate::RY:
                emit_1q(OP_RY, qubits(), OPF_RANDOMIZE);
                P.has_randomizing_ops = true;
                return;
            case Gate::M: measure_1q(in, k, OP_M); return;
            case Gate::MX: measure_1q(in, k, OP_MX); return;
            case Gate::MY: measure_1q(in, k, OP_MY); return;
            case Gate::MR: measure_1q(in, k, OP_MR); return;
            case Gate::MRX: measure_1q(in, k, OP_MRX); return;
            case Gate::MRY: measure_1q(in, k, OP_MRY); return;
            case Gate::MPAD: {
                uint64_t m_first = mcount;
                std::vector<std::pair<uint32_t, uint32_t>> ranges;
                emit_records(OP_RECZERO, T.size(), [&](size_t) {}, nullptr, 0, &ranges);
                if (!in.args.empty()) emit_mflips(in.args[0], m_first, ranges, k);
                return;
            }
            case Gate::MXX: case Gate::MYY: case Gate::MZZ: {
                uint32_t pb = in.gate == Gate::MXX ? TERM_X : in.gate == Gate::MZZ ? TERM_Z : (TERM_X | TERM_Z);
                std::vector<std::vector<uint32_t>> prods;
                for (size_t i = 0; i < T.size(); i += 2) {
                    uint32_t a = q(T[i]), b = q(T[i + 1]);
                    prods.push_back({TERM_PAULI | a | pb, TERM_PAULI | b | pb});
                }
                measure_products(in, k, prods);
                return;
            }
            case Gate::MPP: {
                std::vector<std::vector<uint32_t>> prods;
                for (auto &g : product_groups(in)) prods.push_back(combine_product(in, g.first, g.second, nullptr));
                measure_products(in, k, prods);
                return;
            }
            case Gate::SPP: case Gate::SPP_DAG: {
                for (auto &g : product_groups(in)) {
                    std::vector<uint32_t> bits;
                    std::vector<uint32_t> terms = combine_product(in, g.first, g.second, &bits);
                    if (terms.empty()) continue;
                    std::vector<uint32_t> words = terms;
                    for (uint32_t s : bits) words.push_back(s);
                    std::vector<std::vector<uint32_t>> one{words};
                    emit_csr(
                        OP_SPP, one,
                        [&](size_t) {
                            for (uint32_t t : terms) touch_q(t & TERM_QMASK, KW);
                            for (uint32_t s : bits) touch_r(s, KR);
                        },
                        [&](DevOp &, size_t) {});
                }
                return;
            }
            default:
                break;
        }
        // Two-qubit gates with possible classical controls.
        if (gate_info(in.gate).shape == Shape::PAIRS && !gate_info(in.gate).produces_results &&
            in.gate != Gate::DEPOLARIZE2 && in.gate != Gate::PAULI_CHANNEL_2) {
            uint32_t code = 0;
            bool swap_operands = false;
            switch (in.gate) {
                case Gate::CX: code = OP_CX; break;
                case Gate::CY: code = OP_CY; break;
                case Gate::CZ: code = OP_CZ; break;
                case Gate::XCZ: code = OP_CX; swap_operands = true; break;
                case Gate::YCZ: code = OP_CY; swap_operands = true; break;
                case Gate::SWAP: code = OP_SWAP; break;
                case Gate::ISWAP: case Gate::ISWAP_DAG: code = OP_ISWAP; break;
                case Gate::CXSWAP: code = OP_CXSWAP; break;
                case Gate::CZSWAP: code = OP_CZSWAP; break;
                case Gate::SWAPCX: code = OP_SWAPCX; break;
                case Gate::SQRT_XX: case Gate::SQRT_XX_DAG: code = OP_SQRTXX; break;
                case Gate::SQRT_YY: case Gate::SQRT_YY_DAG: code = OP_SQRTYY; break;
                case Gate::SQRT_ZZ: case Gate::SQRT_ZZ_DAG: code = OP_SQRTZZ; break;
                case Gate::XCX: code = OP_XCX; break;
                case Gate::XCY: code = OP_XCY; break;
                case Gate::YCX: code = OP_YCX; break;
                case Gate::YCY: code = OP_YCY; break;
                default:
                    throw std::invalid_argument(std::string("Not implemented in the frame sampler: ") + gate_info(in.gate).name);
            }
            // Keep instruction order: runs of quantum pairs and runs of classical-control items.
            std::vector<std::pair<uint32_t, uint32_t>> qpairs;
            std::vector<std::array<uint32_t, 3>> fb;
            auto flush_runs = [&]() {
                if (!qpairs.empty()) emit_2q(code, qpairs), qpairs.clear();
                if (!fb.empty()) emit_feedback(fb), fb.clear();
            };
            for (size_t i = 0; i < T.size(); i += 2) {
                Target a = T[i], b = T[i + 1];
                if (swap_operands) std::swap(a, b);
                bool abit = a.kind != TK::QUBIT, bbit = b.kind != TK::QUBIT;
                if (!abit && !bbit) {
                    if (!fb.empty()) flush_runs();
                    qpairs.push_back({q(a), q(b)});
                    continue;
                }
                if (!qpairs.empty()) flush_runs();
                if (code == OP_CZ) {
                    if (abit && bbit) continue;
                    const Target &bit = abit ? a : b;
                    const Target &qq = abit ? b : a;
                    if (bit.kind == TK::REC) fb.push_back({lookback_slot(bit.value), q(qq), 2u});
                    continue;
                }
                // CX / CY: control must be the bit.
                if (bbit) throw std::invalid_argument("Controlled gate had a bit as its target, instead of its control.");
                if (a.kind == TK::REC) fb.push_back({lookback_slot(a.value), q(b), code == OP_CX ? 1u : 3u});
            }
            flush_runs();
            return;
        }
        // Noise channels.
        switch (in.gate) {
            case Gate::X_ERROR: case Gate::Y_ERROR: case Gate::Z_ERROR: case Gate::DEPOLARIZE1:
            case Gate::PAULI_CHANNEL_1: {
                double p;
                std::vector<std::pair<double, uint32_t>> outs;
                if (in.gate == Gate::X_ERROR) p = in.args[0], outs = {{1.0, 1u}};
                else if (in.gate == Gate::Y_ERROR) p = in.args[0], outs = {{1.0, 3u}};
                else if (in.gate == Gate::Z_ERROR) p = in.args[0], outs = {{1.0, 2u}};
                else if (in.gate == Gate::DEPOLARIZE1) p = in.args[0], outs = {{1.0, 1u}, {1.0, 2u}, {1.0, 3u}};
                else {
                    p = in.args[0] + in.args[1] + in.args[2];
                    outs = {{in.args[0], 1u}, {in.args[1], 3u}, {in.args[2], 2u}};
                }
                std::vector<uint32_t> qs = qubits();
                if (qs.empty()) return;
                uint32_t ntoff = (uint32_t)P.ntgt.size();
                for (uint32_t x : qs) {
                    P.ntgt.push_back(x);
                    touch_q(x, KA);
                }
                emit_noise(OP_NOISE1, (uint32_t)qs.size(), ntoff, std::min(p, 1.0), outs, k, 0);
                return;
            }
            case Gate::DEPOLARIZE2: case Gate::PAULI_CHANNEL_2: {
                double p;
                std::vector<std::pair<double, uint32_t>> outs;
                if (in.gate == Gate::DEPOLARIZE2) {
                    p = in.args[0];
                    for (uint32_t c = 1; c < 16; c++) outs.push_back({1.0, c});
                } else {
                    p = 0;
                    const uint32_t m[4] = {0, 1, 3, 2};  // pauli index 1=X,2=Y,3=Z -> (x,z) bits
                    for (uint32_t i = 1; i < 16; i++) {
                        p += in.args[i - 1];
                        outs.push_back({in.args[i - 1], m[(i >> 2) & 3] | (m[i & 3] << 2)});
                    }
                }
                if (T.empty()) return;
                uint32_t ntoff = (uint32_t)P.ntgt.size();
                for (const Target &t : T) {
                    P.ntgt.push_back(q(t));
                    touch_q(q(t), KA);
                }
                emit_noise(OP_NOISE2, (uint32_t)T.size() / 2, ntoff, std::min(p, 1.0), outs, k, 0);
                return;
            }
            case Gate::E: case Gate::ELSE_CORRELATED_ERROR: {
                if (in.gate == Gate::E) {
                    DevOp op = make(OP_CHAINCLR, 1);
                    touch_m(MISC_CHAIN, KW);
                    push(op);
                }
                uint32_t ntoff = (uint32_t)P.ntgt.size();
                for (const Target &t : T) {
                    uint32_t term = TERM_PAULI | q(t) | ((t.pauli & 1) ? TERM_X : 0) | ((t.pauli & 2) ? TERM_Z : 0);
                    P.ntgt.push_back(term);
                    touch_q(q(t), KA);
                }
                touch_m(MISC_CHAIN, KA);
                uint32_t opi = emit_noise(OP_CORR, 1, ntoff, std::min(in.args[0], 1.0), {{1.0, 1u}}, k, 0,
                                          in.gate == Gate::E ? OPF_CHAIN_START : 0);
                P.ops[opi].base = (uint32_t)T.size();  // term count
                return;
            }
            case Gate::HERALDED_ERASE: case Gate::HERALDED_PAULI_CHANNEL_1: {
                std::vector<uint32_t> qs = qubits();
                uint64_t m_first = mcount;
                std::vector<std::pair<uint32_t, uint32_t>> ranges;
                emit_records(OP_RECZERO, qs.size(), [&](size_t) {}, nullptr, 0, &ranges);
                double p;
                std::vector<std::pair<double, uint32_t>> outs;
                if (in.gate == Gate::HERALDED_ERASE) {
                    p = in.args[0];
                    outs = {{1.0, 4u}, {1.0, 5u}, {1.0, 6u}, {1.0, 7u}};
                } else {
                    p = in.args[0] + in.args[1] + in.args[2] + in.args[3];
                    outs = {{in.args[0], 4u}, {in.args[1], 5u}, {in.args[2], 7u}, {in.args[3], 6u}};
                }
                for (const auto &rg : ranges) {
                    uint32_t ntoff = (uint32_t)P.ntgt.size();
                    for (uint32_t j = 0; j < rg.second; j++) {
                        P.ntgt.push_back(qs[rg.first + j]);
                        touch_q(qs[rg.first + j], KA);
                        touch_r(slot_of(m_first + rg.first + j), KA);
                    }
                    uint32_t opi = emit_noise(OP_HERALD, rg.second, ntoff, std::min(p, 1.0), outs, k, rg.first);
                    P.ops[opi].base = slot_of(m_first + rg.first);
                }
                return;
            }
            default:
                throw std::invalid_argument(std::string("Not implemented in the frame sampler: ") + gate_info(in.gate).name);
        }
    }
};

}  // namespace

Program compile_program(const std::vector<const Instr *> &flat, bool measurement_outputs) {
    Builder b;
    b.meas_outputs = measurement_outputs;
    b.run(flat);
    return std::move(b.P);
}

void resolve_fault(const Program &prog, uint32_t k, uint32_t j, uint32_t *op, uint32_t *site) {
    if (k >= prog.sites.size()) {
        throw std::out_of_range("fault instruction index " + std::to_string(k) + " is past the end of the circuit");
    }
    for (const SiteRange &r : prog.sites[k]) {
        if (j >= r.first && j < r.first + r.count) {
            *op = r.op;
            *site = j - r.first;
            return;
        }
    }
    throw std::out_of_range("instruction " + std::to_string(k) + " has no noise site " + std::to_string(j));
}

size_t lds_bytes(const Program &prog, int T) {
    size_t rows = 2 * (size_t)prog.num_qubits + prog.ring + STAGE_ROWS + prog.num_observables + 1;
    return rows * 4 * (size_t)T + 2 * STAGE_WORDS * 4;
}

}  // namespace sa
