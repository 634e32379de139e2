// Planner of the measurement sampler's expansion path (semantics in expand.h).
#include "expand.h"

#include <algorithm>
#include <numeric>

namespace sa {

namespace {

// a <- a XOR b for sorted index sets.
void sym_diff(std::vector<uint32_t> &a, const std::vector<uint32_t> &b, std::vector<uint32_t> &tmp) {
    tmp.clear();
    size_t i = 0, j = 0;
    while (i < a.size() && j < b.size()) {
        if (a[i] < b[j]) {
            tmp.push_back(a[i++]);
        } else if (b[j] < a[i]) {
            tmp.push_back(b[j++]);
        } else {
            i++;
            j++;
        }
    }
    tmp.insert(tmp.end(), a.begin() + i, a.end());
    tmp.insert(tmp.end(), b.begin() + j, b.end());
    a.swap(tmp);
}

// Sorted set of the indices that appear an odd number of times.
std::vector<uint32_t> odd_set(std::vector<uint32_t> v) {
    std::sort(v.begin(), v.end());
    size_t w = 0;
    for (size_t i = 0; i < v.size();) {
        if (i + 1 < v.size() && v[i] == v[i + 1]) {
            i += 2;
        } else {
            v[w++] = v[i++];
        }
    }
    v.resize(w);
    return v;
}

}  // namespace

ExpandPlan plan_expansion(const std::vector<const Instr *> &flat, uint32_t M, uint32_t D, uint32_t O,
                          uint64_t random_recorded, uint64_t random_total, uint32_t threads) {
    ExpandPlan E;
    E.M = M;
    E.Y = D + O;
    E.threads = threads;
    if (M >= (1u << 16)) {
        E.why = "more than 65,535 measurements";
        return E;
    }
    // Rows of A: each detector's measurement set, then each observable's (absolute record indices).
    std::vector<std::vector<uint32_t>> rows;
    std::vector<uint32_t> row_y;
    std::vector<std::vector<uint32_t>> obs_terms(O);
    std::vector<uint8_t> obs_frame(O, 0);
    uint64_t m = 0;
    for (const Instr *in : flat) {
        if (in->gate == Gate::DETECTOR) {
            std::vector<uint32_t> ts;
            for (const Target &t : in->targets) {
                if (t.kind == TK::REC) ts.push_back((uint32_t)(m - t.value));
            }
            row_y.push_back((uint32_t)rows.size());
            rows.push_back(odd_set(std::move(ts)));
        } else if (in->gate == Gate::OBSERVABLE_INCLUDE) {
            const size_t k = (size_t)in->args[0];
            for (const Target &t : in->targets) {
                if (t.kind == TK::REC) obs_terms[k].push_back((uint32_t)(m - t.value));
                else obs_frame[k] = 1;
            }
        }
        m += num_results(*in);
    }
    if (rows.size() != D || m != M) {
        E.why = "record bookkeeping mismatch";
        return E;
    }
    for (uint32_t k = 0; k < O; k++) {
        if (obs_frame[k]) continue;   // not a record parity: no constraint
        row_y.push_back(D + k);
        rows.push_back(odd_set(std::move(obs_terms[k])));
    }

    // Elimination, rows in order of their largest index: a reduced row's pivot is its largest index,
    // and C holds the y columns whose XOR it equals.
    struct Reduced {
        std::vector<uint32_t> R, C;
    };
    std::vector<Reduced> red;
    std::vector<int32_t> pivot_of(M, -1);
    std::vector<uint32_t> order(rows.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const uint32_t ka = rows[a].empty() ? 0 : rows[a].back() + 1, kb = rows[b].empty() ? 0 : rows[b].back() + 1;
        return ka < kb;
    });
    const uint64_t budget = 256ull * ((uint64_t)M + rows.size()) + (1ull << 22);
    uint64_t work = 0;
    std::vector<uint32_t> tmp;
    for (uint32_t r : order) {
        Reduced x{rows[r], {row_y[r]}};
        while (!x.R.empty()) {
            const uint32_t p = x.R.back();
            if (pivot_of[p] < 0) {
                pivot_of[p] = (int32_t)red.size();
                red.push_back(std::move(x));
                break;
            }
            const Reduced &q = red[(size_t)pivot_of[p]];
            work += x.R.size() + q.R.size() + x.C.size() + q.C.size();
            if (work > budget) {
                E.why = "elimination of the detectors fills in beyond its budget";
                return E;
            }
            sym_diff(x.R, q.R, tmp);
            sym_diff(x.C, q.C, tmp);
        }
        // A row that reduced to nothing is a combination of earlier rows: its y follows from theirs.
    }
    E.pivots = (uint32_t)red.size();
    const uint64_t nfree = (uint64_t)M - red.size();
    (void)random_total;
    if (nfree != random_recorded) {
        E.why = "the detectors and observables do not fix every parity the noiseless circuit fixes (" +
                std::to_string(nfree) + " unconstrained results, " + std::to_string(random_recorded) + " random ones)";
        return E;
    }

    // y contributions: each y column's first slot inline, further slots in a side list.
    std::vector<std::vector<uint32_t>> slots_of(E.Y);
    for (const Reduced &x : red) {
        for (uint32_t y : x.C) slots_of[y].push_back(x.R.back());
    }
    E.prim.assign(E.Y, 0xFFFFFFFFu);
    E.extra_off.assign(E.Y + 1, 0);
    for (uint32_t y = 0; y < E.Y; y++) {
        const auto &s = slots_of[y];
        if (!s.empty()) E.prim[y] = s[0] | (s.size() > 1 ? 0x80000000u : 0u);
        for (size_t i = 1; i < s.size(); i++) E.extra.push_back(s[i]);
        E.extra_off[y + 1] = (uint32_t)E.extra.size();
    }
    // Free results; items with their owner thread and phase (record order is a topological order).
    std::vector<int32_t> owner(M, -1);
    std::vector<uint32_t> phase(M, 0), level(M, 0);
    uint32_t next_thread = 0;
    std::vector<uint32_t> item_list;
    for (uint32_t k = 0; k < M; k++) {
        if (pivot_of[k] < 0) {
            E.free_slots.push_back(k);
            continue;
        }
        const Reduced &x = red[(size_t)pivot_of[k]];
        if (x.R.size() < 2) continue;   // no earlier terms: its value is in place after the y pass
        int32_t own = -1;
        uint32_t best = 0, lv = 0;
        for (size_t i = 0; i + 1 < x.R.size(); i++) {
            const uint32_t t = x.R[i];
            lv = std::max(lv, level[t] + 1);
            if (owner[t] >= 0 && (own < 0 || phase[t] > best)) {
                own = owner[t];
                best = phase[t];
            }
        }
        if (own < 0) own = (int32_t)(next_thread++ % threads);
        uint32_t ph = 0;
        for (size_t i = 0; i + 1 < x.R.size(); i++) {
            const uint32_t t = x.R[i];
            if (owner[t] >= 0) ph = std::max(ph, phase[t] + (owner[t] != own ? 1u : 0u));
        }
        owner[k] = own;
        phase[k] = ph;
        level[k] = lv;
        E.levels = std::max(E.levels, lv);
        E.phases = std::max(E.phases, ph + 1);
        item_list.push_back(k);
    }
    // Items grouped by (phase, thread), record order within a group.
    std::stable_sort(item_list.begin(), item_list.end(), [&](uint32_t a, uint32_t b) {
        if (phase[a] != phase[b]) return phase[a] < phase[b];
        return owner[a] < owner[b];
    });
    // Records. Consecutive items of one (phase, thread) whose terms include the previous item's
    // result, at a constant record distance, form a run: each result is the previous one XOR its own
    // value (its y contributions and its other terms). Where the later items have no other terms the
    // run is one chain record. Where their other terms (and the first item's terms, when it has
    // several) are final when the phase starts (free results, results without substitution, results
    // of earlier phases), those terms are XORed into the run's results by independent records in a
    // phase of their own just before, spread over the threads, and the run is again one chain record:
    // its thread then walks the run reading 8 results per LDS round trip instead of one record per
    // round trip. Otherwise each item is a record of its own.
    auto final_at = [&](uint32_t t, uint32_t P) { return owner[t] < 0 || phase[t] < P; };
    struct Rec {
        uint32_t phase;    // the items' phase
        bool pre;          // a pre-phase record (XORs final terms into a run's result)
        uint32_t thread;
        uint32_t w[EXPAND_RECORD_WORDS];
        std::vector<uint32_t> ovf;   // terms of a record with more than EXPAND_INLINE_TERMS
    };
    std::vector<Rec> recs;
    std::vector<uint8_t> has_pre(E.phases, 0);
    auto add_record = [&](uint32_t P, bool pre, uint32_t thread, uint32_t k, const std::vector<uint32_t> &ts) {
        const uint32_t n = (uint32_t)ts.size();
        Rec r{P, pre, thread, {k | (n << 16), 0, 0, 0, 0, 0, 0, 0}, {}};
        if (n == 1) {
            r.w[1] = ts[0];   // a chain of one
            r.w[2] = 1;
        } else if (n <= EXPAND_INLINE_TERMS) {
            for (uint32_t t = 0; t < n; t++) r.w[1 + t] = ts[t];
        } else {
            r.ovf = ts;
        }
        recs.push_back(std::move(r));
    };
    size_t i = 0;
    while (i < item_list.size()) {
        const uint32_t k = item_list[i];
        const Reduced &x = red[(size_t)pivot_of[k]];
        const std::vector<uint32_t> ts(x.R.begin(), x.R.end() - 1);
        if (ts.size() >= (1u << 16)) {
            E.why = "a substitution with more than 65,535 terms";
            return E;
        }
        const uint32_t P = phase[k], th = (uint32_t)owner[k];
        const bool first_final = std::all_of(ts.begin(), ts.end(), [&](uint32_t t) { return final_at(t, P); });
        uint32_t len = 1, stride = 0;
        bool side = false;
        while (i + len < item_list.size()) {
            const uint32_t k2 = item_list[i + len], prev = item_list[i + len - 1];
            if (phase[k2] != P || (uint32_t)owner[k2] != th) break;
            const Reduced &x2 = red[(size_t)pivot_of[k2]];
            if (!std::binary_search(x2.R.begin(), x2.R.end() - 1, prev)) break;
            if (len == 1) stride = k2 - prev;
            else if (k2 - prev != stride) break;
            if (x2.R.size() > 2) {
                bool ok = ts.size() == 1 || first_final;
                for (size_t t = 0; ok && t + 1 < x2.R.size(); t++) ok = x2.R[t] == prev || final_at(x2.R[t], P);
                if (!ok) break;
                side = true;
            }
            len++;
        }
        if (len > 1 && ts.size() > 1 && !first_final) len = 1;   // side is false here: no pre-record for it
        if (len == 1) {
            add_record(P, false, th, k, ts);
            i++;
            continue;
        }
        uint32_t y0 = ts[0];
        if (ts.size() > 1) {
            add_record(P, true, 0, k, ts);
            y0 = M;   // a result that is always zero
            has_pre[P] = 1;
        }
        for (uint32_t j = 1; side && j < len; j++) {
            const uint32_t kj = item_list[i + j], prev = item_list[i + j - 1];
            const Reduced &xj = red[(size_t)pivot_of[kj]];
            std::vector<uint32_t> other;
            for (size_t t = 0; t + 1 < xj.R.size(); t++) {
                if (xj.R[t] != prev) other.push_back(xj.R[t]);
            }
            if (!other.empty()) add_record(P, true, 0, kj, other);
            has_pre[P] = 1;
        }
        Rec c{P, false, th, {k | (1u << 16), y0, len, stride, 0, 0, 0, 0}, {}};
        recs.push_back(std::move(c));
        i += len;
    }
    // Phases with pre-records get a phase of their own just before them; pre-records go round the
    // threads of that phase.
    std::vector<uint32_t> newp(E.phases, 0);
    uint32_t shift = 0;
    for (uint32_t P = 0; P < E.phases; P++) {
        shift += has_pre[P];
        newp[P] = P + shift;
    }
    E.phases += shift;
    std::vector<uint32_t> rr(E.phases, 0);
    for (Rec &r : recs) {
        const uint32_t ph = newp[r.phase] - (r.pre ? 1u : 0u);
        if (r.pre) r.thread = rr[ph]++ % threads;
        r.phase = ph;
    }
    std::stable_sort(recs.begin(), recs.end(), [](const Rec &a, const Rec &b) {
        return a.phase != b.phase ? a.phase < b.phase : a.thread < b.thread;
    });
    E.sched_off.assign((size_t)E.phases * threads + 1, 0);
    for (Rec &r : recs) {
        if (!r.ovf.empty()) {
            r.w[1] = (uint32_t)E.terms.size();
            E.terms.insert(E.terms.end(), r.ovf.begin(), r.ovf.end());
        }
        E.items.insert(E.items.end(), r.w, r.w + EXPAND_RECORD_WORDS);
        E.sched_off[(size_t)r.phase * threads + r.thread + 1]++;
    }
    for (size_t g = 1; g < E.sched_off.size(); g++) E.sched_off[g] += E.sched_off[g - 1];
    E.complete = true;
    return E;
}

void expand_shot_host(const ExpandPlan &E, const uint8_t *y, const uint8_t *free_bits, uint8_t *m) {
    std::fill(m, m + E.M, 0);
    for (size_t f = 0; f < E.free_slots.size(); f++) m[E.free_slots[f]] = free_bits[f] & 1;
    for (uint32_t j = 0; j < E.Y; j++) {
        if (!(y[j] & 1) || E.prim[j] == 0xFFFFFFFFu) continue;
        m[E.prim[j] & 0xFFFFu] ^= 1;
        if (E.prim[j] >> 31) {
            for (uint32_t c = E.extra_off[j]; c < E.extra_off[j + 1]; c++) m[E.extra[c]] ^= 1;
        }
    }
    // Phases in order, threads in order within a phase: a topological order of the records.
    const size_t nrec = E.sched_off.empty() ? 0 : E.sched_off.back();
    for (size_t i = 0; i < nrec; i++) {
        const uint32_t *r = &E.items[EXPAND_RECORD_WORDS * i];
        uint32_t k = r[0] & 0xFFFFu;
        const uint32_t n = r[0] >> 16;
        uint8_t v = m[k];
        if (n == 1) {
            v ^= r[1] < E.M ? m[r[1]] : 0;   // index M: the always-zero result
            m[k] = v;
            for (uint32_t j = 1; j < r[2]; j++) {
                k += r[3];
                v ^= m[k];
                m[k] = v;
            }
        } else if (n <= EXPAND_INLINE_TERMS) {
            for (uint32_t t = 0; t < n; t++) v ^= m[r[1 + t]];
            m[k] = v;
        } else {
            for (uint32_t t = 0; t < n; t++) v ^= m[E.terms[r[1] + t]];
            m[k] = v;
        }
    }
}

}  // namespace sa
