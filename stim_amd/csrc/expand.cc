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
                          uint64_t random_recorded, uint64_t random_total) {
    ExpandPlan E;
    E.M = M;
    E.Y = D + O;
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

    // y contributions (CSR over y columns).
    std::vector<uint32_t> count(E.Y + 1, 0);
    for (const Reduced &x : red) {
        for (uint32_t y : x.C) count[y + 1]++;
    }
    E.contrib_off.assign(E.Y + 1, 0);
    for (uint32_t y = 0; y < E.Y; y++) E.contrib_off[y + 1] = E.contrib_off[y] + count[y + 1];
    E.contrib.assign(E.contrib_off[E.Y], 0);
    std::vector<uint32_t> fill(E.contrib_off.begin(), E.contrib_off.end() - 1);
    for (const Reduced &x : red) {
        for (uint32_t y : x.C) E.contrib[fill[y]++] = x.R.back();
    }
    // Free results and substitution levels (record order is a topological order).
    std::vector<uint32_t> level(M, 0);
    uint32_t levels = 0;
    for (uint32_t k = 0; k < M; k++) {
        if (pivot_of[k] < 0) {
            E.free_slots.push_back(k);
            continue;
        }
        const Reduced &x = red[(size_t)pivot_of[k]];
        uint32_t l = 0;
        for (size_t i = 0; i + 1 < x.R.size(); i++) l = std::max(l, level[x.R[i]] + 1);
        level[k] = l;
        levels = std::max(levels, l);
    }
    std::vector<std::vector<uint32_t>> by_level(levels + 1);
    for (uint32_t k = 0; k < M; k++) {
        if (pivot_of[k] >= 0 && level[k] > 0) by_level[level[k]].push_back(k);
    }
    E.level_off.push_back(0);
    E.term_off.push_back(0);
    for (uint32_t l = 1; l <= levels; l++) {
        for (uint32_t k : by_level[l]) {
            const Reduced &x = red[(size_t)pivot_of[k]];
            E.item_slot.push_back(k);
            E.terms.insert(E.terms.end(), x.R.begin(), x.R.end() - 1);
            E.term_off.push_back((uint32_t)E.terms.size());
        }
        E.level_off.push_back((uint32_t)E.item_slot.size());
    }
    E.complete = true;
    return E;
}

void expand_shot_host(const ExpandPlan &E, const uint8_t *y, const uint8_t *free_bits, uint8_t *m) {
    std::fill(m, m + E.M, 0);
    for (size_t f = 0; f < E.free_slots.size(); f++) m[E.free_slots[f]] = free_bits[f] & 1;
    for (uint32_t j = 0; j < E.Y; j++) {
        if (!(y[j] & 1)) continue;
        for (uint32_t c = E.contrib_off[j]; c < E.contrib_off[j + 1]; c++) m[E.contrib[c]] ^= 1;
    }
    for (size_t i = 0; i < E.item_slot.size(); i++) {
        uint8_t v = m[E.item_slot[i]];
        for (uint32_t t = E.term_off[i]; t < E.term_off[i + 1]; t++) v ^= m[E.terms[t]];
        m[E.item_slot[i]] = v;
    }
}

}  // namespace sa
