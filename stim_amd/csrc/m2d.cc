// Measurement-to-detection-event converter behind the C ABI (mfs_m2d_*).
//
// Mirrors the reference's CompiledMeasurementsToDetectionEventsConverter
// (src/stim/simulators/measurements_to_detection_events.pybind.cc:78-159) without sweep data: the
// circuit's detectors and observables become column records (up to two absolute measurement indices
// inline, longer lists in an overflow table) with the parity of their measurements' reference-sample
// bits (semantics and layout in m2d.h). Conversion runs on
// the GPU (m2d_kernel.hip); host rows are staged through device buffers with 16-byte-aligned strides.
#include "m2d.h"

#include "sampler_impl.h"
#include "tableau.h"

namespace sa {
hipError_t launch_m2d(const M2dParams &P, hipStream_t stream);
size_t m2d_lds_bytes(uint32_t blocks, uint32_t sweep_blocks);
hipError_t launch_m2d_win(const M2dParams &P, const M2dWindow &Wn, hipStream_t stream);
size_t m2d_win_lds_bytes(uint32_t ring_words, uint32_t nchunks);
}  // namespace sa

namespace {

// Number of sweep bits a circuit reads: the largest sweep[k] index plus one (the reference's
// CircuitStats::num_sweep_bits).
uint32_t count_sweep_bits(const std::vector<const Instr *> &flat) {
    uint64_t k = 0;
    for (const Instr *in : flat) {
        for (const Target &t : in->targets) {
            if (t.kind == TK::SWEEP) k = std::max<uint64_t>(k, (uint64_t)t.value + 1);
        }
    }
    if (k >= (1ull << 24)) throw std::invalid_argument("too many sweep bits for the converter");
    return (uint32_t)k;
}

// Copies a block, replacing each sweep-controlled pair of CX/CY/CZ/XCZ/YCZ by a one-qubit error
// instruction on its qubit, tagged k + 1 for sweep bit k: the frame flip the reference's frame
// simulator applies when that sweep bit is set (xor_control_bit_into, frame_simulator.inl:139-150:
// CX and XCZ flip X, CY and YCZ flip X and Z, CZ flips Z; single_cx/single_cy/do_ZCZ at :386-474,
// do_XCZ/do_YCZ at :592-630 with their pairs reversed). Pairs of two classical bits do nothing; the
// other pairs keep their order, split into instructions of the same gate.
void split_sweep_controls(const Block &src, Block &dst) {
    for (const Item &it : src.items) {
        if (it.is_repeat) {
            Item r;
            r.is_repeat = true;
            r.reps = it.reps;
            r.body = std::make_shared<Block>();
            split_sweep_controls(*it.body, *r.body);
            dst.items.push_back(std::move(r));
            continue;
        }
        const Instr &in = it.instr;
        const Gate g = in.gate;
        bool has_sweep = false;
        for (const Target &t : in.targets) has_sweep |= t.kind == TK::SWEEP;
        const bool controllable = g == Gate::CX || g == Gate::CY || g == Gate::CZ || g == Gate::XCZ || g == Gate::YCZ;
        if (!has_sweep || !controllable) {
            dst.items.push_back(it);
            continue;
        }
        Instr pending{g, in.args, {}, 0};
        auto flush = [&]() {
            if (pending.targets.empty()) return;
            Item x;
            x.instr = pending;
            dst.items.push_back(std::move(x));
            pending.targets.clear();
        };
        for (size_t i = 0; i + 1 < in.targets.size(); i += 2) {
            const Target &a = in.targets[i], &b = in.targets[i + 1];
            if (a.kind != TK::SWEEP && b.kind != TK::SWEEP) {
                pending.targets.push_back(a);
                pending.targets.push_back(b);
                continue;
            }
            // Control and target of the pair in the reference's single_cx / single_cy orientation.
            const bool reversed = g == Gate::XCZ || g == Gate::YCZ;
            const Target &ctl = reversed ? b : a, &tgt = reversed ? a : b;
            Gate err;
            Target q;
            if (g == Gate::CZ) {
                if (a.kind == TK::SWEEP && b.kind == TK::QUBIT) q = b;
                else if (b.kind == TK::SWEEP && a.kind == TK::QUBIT) q = a;
                else continue;
                err = Gate::Z_ERROR;
            } else {
                if (ctl.kind != TK::SWEEP || tgt.kind != TK::QUBIT) {
                    if (tgt.kind != TK::QUBIT) throw std::invalid_argument("a controlled gate had a bit as its target");
                    continue;
                }
                q = tgt;
                err = (g == Gate::CX || g == Gate::XCZ) ? Gate::X_ERROR : Gate::Y_ERROR;
            }
            const uint32_t k = (ctl.kind == TK::SWEEP) ? ctl.value : (a.kind == TK::SWEEP ? a.value : b.value);
            flush();
            Item x;
            x.instr = Instr{err, {0.5}, {q}, k + 1};
            dst.items.push_back(std::move(x));
        }
        flush();
    }
}

// Effect of each sweep bit on the circuit's outputs: rows[k] = the detectors then observables (bit
// packed, little-endian, D + O bits) that sweep bit k flips. Runs the frame kernel in fixed-fault
// mode on the circuit with its sweep controls turned into error sites, one shot per sweep bit with
// every site of that bit fired (no frame randomization, no other noise: the reference's noiseless,
// non-randomizing frame simulator of measurements_to_detection_events.inl:57-63).
std::vector<uint8_t> sweep_effect_rows(const Block &block, uint32_t K, uint32_t D, uint32_t O, size_t row) {
    auto s = std::make_unique<mfs_sampler>();
    split_sweep_controls(block, s->block);
    s->flat = flatten(s->block);
    s->prog = compile_program(s->flat);
    s->key = 0;
    std::vector<std::vector<uint32_t>> per_bit(K);
    for (size_t i = 0; i < s->flat.size(); i++) {
        const Instr *in = s->flat[i];
        if (!in->tag) continue;
        const uint32_t code = in->gate == Gate::X_ERROR ? 1u : in->gate == Gate::Z_ERROR ? 2u : 3u;
        per_bit[in->tag - 1].insert(per_bit[in->tag - 1].end(), {(uint32_t)i, 0u, code});
    }
    std::vector<uint64_t> offs(K + 1, 0);
    std::vector<uint32_t> faults;
    for (uint32_t k = 0; k < K; k++) {
        faults.insert(faults.end(), per_bit[k].begin(), per_bit[k].end());
        offs[k + 1] = faults.size() / 3;
    }
    if (faults.empty()) return std::vector<uint8_t>((size_t)K * row, 0);
    check(hipStreamCreate(&s->stream), "stream");
    s->pick_tile_words(0);
    s->upload_program();
    s->engine_mode = 1;
    std::vector<uint8_t> rows((size_t)K * row, 0);
    // No error site at all (sweep bits that only meet classical bits): every effect is empty.
    if (D + O && !faults.empty())
        s->sample(K, MFS_BIT_PACKED | MFS_APPEND_OBSERVABLES, rows.data(), row, nullptr, 0, offs.data(), faults.data());
    return rows;
}

// Largest ring of the windowed kernel (words): 32 KB of LDS, with the staging image about 41 KB per
// 256-lane workgroup.
constexpr uint32_t MAX_RING_WORDS = 8192;

// Window of one output layout: part 0 = columns [0, n0), emitted block by block as their terms arrive;
// with `tail`, the observable columns [D, cols) also go to separate rows after the last chunk.
// ring_words = 0 when some column's terms lie further apart than MAX_RING_WORDS allows.
struct WinPlan {
    uint32_t ring_words = 0;
    std::vector<uint32_t> chunk_end;   // per chunk: part-0 blocks whose last term it holds, cumulative
};

WinPlan plan_window(const std::vector<std::vector<uint32_t>> &col_terms, uint32_t n0, uint32_t D, bool tail,
                    uint32_t M) {
    WinPlan W;
    const uint32_t nchunks = (M + M2D_WIN_CHUNK - 1) / M2D_WIN_CHUNK;
    uint64_t need = M2D_WIN_CHUNK;   // a ring holds at least one chunk
    uint32_t prev = 0;
    W.chunk_end.assign(std::max<uint32_t>(nchunks, 1), 0);
    for (uint32_t b = 0; b * 32 < n0; b++) {
        uint32_t hi = 0;
        for (uint32_t j = 32 * b; j < std::min(n0, 32 * b + 32); j++) {
            for (uint32_t t : col_terms[j]) hi = std::max(hi, t / M2D_WIN_CHUNK);
        }
        prev = std::max(prev, hi);
        W.chunk_end[prev] = b + 1;
        // After chunk `prev` is loaded the ring holds results [(prev + 1) * CHUNK - ring, (prev + 1) * CHUNK).
        for (uint32_t j = 32 * b; j < std::min(n0, 32 * b + 32); j++) {
            for (uint32_t t : col_terms[j]) need = std::max<uint64_t>(need, (uint64_t)(prev + 1) * M2D_WIN_CHUNK - t);
        }
    }
    for (size_t j = D; tail && j < col_terms.size(); j++) {
        for (uint32_t t : col_terms[j]) need = std::max<uint64_t>(need, (uint64_t)nchunks * M2D_WIN_CHUNK - t);
    }
    uint32_t ring = M2D_WIN_CHUNK;
    while (ring < need && ring <= MAX_RING_WORDS / 2) ring *= 2;
    if (ring >= need && nchunks <= M2D_WIN_MAX_CHUNKS) W.ring_words = ring;
    for (size_t ch = 1; ch < W.chunk_end.size(); ch++) W.chunk_end[ch] = std::max(W.chunk_end[ch], W.chunk_end[ch - 1]);
    return W;
}

}  // namespace

struct mfs_m2d {
    uint32_t M = 0, D = 0, O = 0;
    uint32_t K = 0;   // sweep bits the circuit reads
    std::vector<uint2> cols;
    std::vector<uint32_t> ovf;
    DevBuf d_cols, d_ovf, d_in, d_out, d_obs, d_tmp, d_sweep;
    // Windowed kernel (no sweep bits), one plan per output layout: 0 = observables appended, 1 =
    // observables separate, 2 = detectors only, 3 = observables appended and separate. kernel_mode:
    // 0 = the whole-record kernel where the tile's record fits in LDS, else windowed; 1 = whole-record
    // kernel; 2 = windowed (refused where it does not apply).
    WinPlan win[4];
    DevBuf d_chunk_end[4];
    bool any_window() const { return win[0].ring_words || win[1].ring_words || win[2].ring_words || win[3].ring_words; }
    int kernel_mode = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ~mfs_m2d() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    uint32_t blocks() const { return (M + 31) / 32; }
    uint32_t sweep_blocks() const { return (K + 31) / 32; }
    size_t sweep_row() const { return (K + 7) / 8; }
    size_t in_row() const { return (M + 7) / 8; }
    size_t in_vec_stride() const { return (size_t)16 * ((blocks() + 3) / 4); }

    M2dParams params(uint64_t shots, const uint8_t *in, size_t in_stride, uint8_t *out, size_t out_stride,
                     uint32_t n_out, uint8_t *obs, size_t obs_stride, uint32_t n_obs) const {
        M2dParams P{};
        P.in = in;
        P.in_stride = in_stride;
        P.in_row_bytes = (uint32_t)in_row();
        P.in_vec = ((uintptr_t)in % 16 == 0 && in_stride % 16 == 0 && in_stride >= in_vec_stride()) ? 1u : 0u;
        P.blocks = blocks();
        P.num_det = D;
        P.out = out;
        P.out_stride = out_stride;
        P.n_out = n_out;
        P.n_obs = n_obs;
        P.obs = obs;
        P.obs_stride = obs_stride;
        P.shots = shots;
        P.cols = const_cast<DevBuf &>(d_cols).as<uint2>();
        P.ovf = const_cast<DevBuf &>(d_ovf).as<uint32_t>();
        P.sweep = nullptr;
        P.sweep_stride = 0;
        P.sweep_row_bytes = (uint32_t)sweep_row();
        P.sweep_blocks = sweep_blocks();
        return P;
    }

    // Windowed layout usable for a launch, or -1.
    int window_for(const M2dParams &P) const {
        if (K || !P.in_vec || !M) return -1;
        int l;
        if (P.n_out == D + O) l = P.n_obs ? 3 : 0;
        else if (P.n_out == D) l = P.n_obs ? 1 : 2;
        else return -1;
        return win[l].ring_words ? l : -1;
    }
    bool whole_fits() const { return m2d_lds_bytes(blocks(), sweep_blocks()) <= MAX_LDS; }
    bool use_window(const M2dParams &P) const {
        if (kernel_mode == 1 || window_for(P) < 0) return false;
        return kernel_mode == 2 || !whole_fits();
    }

    void launch(const M2dParams &P) {
        if ((P.shots + 31) / 32 >= (1ull << 31)) throw std::invalid_argument("too many shots in one conversion");
        // The kernel addresses a tile's output rows with 32-bit offsets.
        if (P.out_stride >= (1ull << 30) || P.obs_stride >= (1ull << 30))
            throw std::invalid_argument("output row strides must be below 2^30 bytes");
        if ((kernel_mode == 2 || !whole_fits()) && !P.in_vec && M && !K)
            throw std::invalid_argument("the windowed conversion kernel reads device measurement rows in place only "
                                        "when they are 16-byte aligned with a stride of at least 16 * ceil(M / 128) "
                                        "bytes that is a multiple of 16");
        if ((kernel_mode == 2 || !whole_fits()) && window_for(P) < 0)
            throw std::invalid_argument(kernel_mode == 2
                ? "the windowed conversion kernel does not apply to this circuit or layout"
                : "the measurement record does not fit in LDS, and the windowed conversion kernel does not apply "
                  "to this output layout (observables must be appended, or their measurements must come last)");
        if (use_window(P)) {
            const int l = window_for(P);
            M2dWindow Wn{win[l].ring_words, (M + M2D_WIN_CHUNK - 1) / M2D_WIN_CHUNK,
                         const_cast<DevBuf &>(d_chunk_end[l]).as<uint32_t>(), D + O};
            check(launch_m2d_win(P, Wn, stream), "m2d launch");
            return;
        }
        check(launch_m2d(P, stream), "m2d launch");
    }
};

extern "C" {

int mfs_m2d_compile(const char *stim_text, size_t len, int skip_reference_sample, mfs_m2d **out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = nullptr;
        std::string text(stim_text ? stim_text : "", stim_text ? len : 0);
        Block b = parse_circuit(text.data(), text.size());
        const std::vector<const Instr *> flat = flatten(b);
        const Stats st = compute_stats(flat);
        if (st.num_measurements >= (1ull << 32) || st.num_detectors + st.num_observables >= (1ull << 31))
            throw std::invalid_argument("circuit too large for the converter");
        auto c = std::make_unique<mfs_m2d>();
        c->M = (uint32_t)st.num_measurements;
        c->D = (uint32_t)st.num_detectors;
        c->O = (uint32_t)st.num_observables;
        c->K = count_sweep_bits(flat);
        std::vector<uint8_t> ref(c->M, 0);
        if (!skip_reference_sample) ref = reference_sample(flat);
        // Columns: detectors in order, then observables; terms are absolute measurement indices.
        std::vector<std::vector<uint32_t>> col_terms;
        std::vector<std::vector<uint32_t>> obs_terms(c->O);
        uint64_t m = 0;
        for (const Instr *in : flat) {
            if (in->gate == Gate::DETECTOR) {
                std::vector<uint32_t> ts;
                for (const Target &t : in->targets) ts.push_back((uint32_t)(m - t.value));
                col_terms.push_back(std::move(ts));
            } else if (in->gate == Gate::OBSERVABLE_INCLUDE) {
                for (const Target &t : in->targets) {
                    if (t.kind == TK::REC) obs_terms[(size_t)in->args[0]].push_back((uint32_t)(m - t.value));
                }
            }
            m += num_results(*in);
        }
        for (auto &ts : obs_terms) col_terms.push_back(std::move(ts));
        if (!c->K) {
            c->win[0] = plan_window(col_terms, c->D + c->O, c->D, false, c->M);
            c->win[1] = plan_window(col_terms, c->D, c->D, true, c->M);
            c->win[2] = plan_window(col_terms, c->D, c->D, false, c->M);
            c->win[3] = plan_window(col_terms, c->D + c->O, c->D, true, c->M);
        }
        // The whole-record kernel keeps a tile's measurements and sweep bits in LDS; the windowed one
        // only a ring of recent results.
        if (!c->whole_fits() && !c->any_window()) {
            uint32_t max_blocks = 0;
            while (m2d_lds_bytes(max_blocks + 1, c->sweep_blocks()) <= MAX_LDS) max_blocks++;
            throw std::invalid_argument("the converter keeps a tile's measurements and sweep bits in LDS: at most " +
                                        std::to_string((uint64_t)max_blocks * 32) + " measurements per shot with " +
                                        std::to_string(c->K) + " sweep bits, unless every detector reaches back "
                                        "at most about 7,000 results and the circuit has no sweep bits");
        }
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            throw HipError("no HIP device available: the stim_amd converter runs only on the GPU");
        }
        // The reference parity covers measurement terms only; sweep terms (word indices past the
        // measurements) are appended after it is formed.
        std::vector<uint8_t> invert(col_terms.size(), 0);
        for (size_t j = 0; j < col_terms.size(); j++) {
            for (uint32_t idx : col_terms[j]) invert[j] ^= ref[idx];
        }
        if (c->K) {
            const size_t row = (c->D + c->O + 7) / 8;
            const std::vector<uint8_t> eff = sweep_effect_rows(b, c->K, c->D, c->O, std::max<size_t>(row, 1));
            const uint32_t w0 = m2d_sweep_word(c->blocks());
            for (uint32_t k = 0; k < c->K; k++) {
                const uint8_t *r = eff.data() + (size_t)k * std::max<size_t>(row, 1);
                for (uint32_t j = 0; j < c->D + c->O; j++) {
                    if ((r[j >> 3] >> (j & 7)) & 1) col_terms[j].push_back(w0 + k);
                }
            }
        }
        const uint32_t zero_word = m2d_zero_word(c->blocks(), c->sweep_blocks());
        for (size_t j = 0; j < col_terms.size(); j++) {
            const auto &ts = col_terms[j];
            const uint8_t f = invert[j];
            uint2 r;
            if (ts.size() <= 2) {
                r = make_uint2(((uint32_t)ts.size() << 28) | (ts.size() >= 1 ? ts[0] : zero_word),
                               ts.size() == 2 ? ts[1] : zero_word);
            } else {
                if (ts.size() > M2D_INDEX_MASK || c->ovf.size() + ts.size() + 3 >= (1ull << 32))
                    throw std::invalid_argument("too many measurement terms for the converter");
                r = make_uint2((3u << 28) | (uint32_t)ts.size(), (uint32_t)c->ovf.size());
                c->ovf.insert(c->ovf.end(), ts.begin(), ts.end());
                while (c->ovf.size() % 4) c->ovf.push_back(zero_word);
            }
            if (f) r.x |= M2D_INVERT;
            c->cols.push_back(r);
        }
        // A zero record after the columns (no terms, no inversion): the windowed kernel loads it for
        // lanes past the last column, so no loaded record needs a select.
        c->cols.push_back(make_uint2(0, 0));
        check(hipStreamCreate(&c->stream), "stream");
        mfs_sampler::upload(c->d_cols, c->cols, "m2d columns");
        mfs_sampler::upload(c->d_ovf, c->ovf, "m2d overflow terms");
        for (int l = 0; l < 4; l++) {
            if (c->win[l].ring_words) mfs_sampler::upload(c->d_chunk_end[l], c->win[l].chunk_end, "m2d window");
        }
        *out = c.release();
    });
}

void mfs_m2d_destroy(mfs_m2d *c) {
    delete c;
}

int mfs_m2d_stats(const mfs_m2d *c, uint64_t out[4]) {
    return guarded([&] {
        out[0] = c->M;
        out[1] = c->D;
        out[2] = c->O;
        out[3] = c->K;
    });
}

int mfs_m2d_set_kernel(mfs_m2d *c, int kernel) {
    return guarded([&] {
        if (kernel < 0 || kernel > 2) throw std::invalid_argument("kernel must be 0 (auto), 1 (whole record) or 2 (windowed)");
        if (kernel == 2 && !c->any_window())
            throw std::invalid_argument("the windowed conversion kernel does not apply to this circuit");
        if (kernel == 1 && !c->whole_fits())
            throw std::invalid_argument("the whole-record conversion kernel does not fit this circuit's record in LDS");
        c->kernel_mode = kernel;
    });
}

int mfs_m2d_kernel_info(const mfs_m2d *c, uint64_t out[6]) {
    return guarded([&] {
        for (int l = 0; l < 4; l++) out[l] = c->win[l].ring_words;
        out[4] = m2d_lds_bytes(c->blocks(), c->sweep_blocks());
        out[5] = c->win[0].ring_words ? m2d_win_lds_bytes(c->win[0].ring_words, (c->M + M2D_WIN_CHUNK - 1) / M2D_WIN_CHUNK) : 0;
    });
}

int mfs_m2d_convert(mfs_m2d *c, uint64_t shots, const uint8_t *meas, size_t meas_stride, uint32_t flags,
                    uint8_t *dets, size_t dets_stride, uint8_t *obs, size_t obs_stride) {
    return mfs_m2d_convert_sweep(c, shots, meas, meas_stride, nullptr, 0, flags, dets, dets_stride, obs, obs_stride);
}

int mfs_m2d_convert_sweep(mfs_m2d *c, uint64_t shots, const uint8_t *meas, size_t meas_stride, const uint8_t *sweep,
                          size_t sweep_stride, uint32_t flags, uint8_t *dets, size_t dets_stride, uint8_t *obs,
                          size_t obs_stride) {
    return guarded([&] {
        if (sweep && c->K && sweep_stride < c->sweep_row()) throw std::invalid_argument("sweep bit rows are too short");
        if (!c->K) sweep = nullptr;
        const bool append = flags & MFS_APPEND_OBSERVABLES, separate = flags & MFS_SEPARATE_OBSERVABLES;
        const bool packed = flags & MFS_BIT_PACKED, device = flags & MFS_DEVICE_OUTPUT;
        if (flags & MFS_PREPEND_OBSERVABLES) throw std::invalid_argument("the converter does not prepend observables");
        if (shots == 0) return;
        const uint32_t n_out = c->D + (append ? c->O : 0), n_obs = separate ? c->O : 0;
        const size_t out_row = (n_out + 7) / 8, obs_row = (n_obs + 7) / 8;
        if (c->M && (!meas || meas_stride < c->in_row())) throw std::invalid_argument("measurement rows are too short");
        if (n_out && (!dets || dets_stride < (packed ? out_row : n_out))) throw std::invalid_argument("detection event rows are too short");
        if (n_obs && (!obs || obs_stride < (packed ? obs_row : n_obs))) throw std::invalid_argument("observable rows are too short");
        if (device && packed) {
            M2dParams P = c->params(shots, meas, meas_stride, dets, dets_stride, n_out, obs, obs_stride, n_obs);
            P.sweep = sweep;
            P.sweep_stride = sweep_stride;
            c->launch(P);
            check(hipStreamSynchronize(c->stream), "m2d");
            return;
        }
        // Staged: batches of shots through device buffers (input rows at the vector stride).
        const size_t vs = std::max<size_t>(c->in_vec_stride(), 16);
        const size_t srow = sweep ? c->sweep_row() : 0;
        const size_t per_shot = vs + srow + out_row + obs_row + (packed ? 0 : n_out + n_obs);
        const uint64_t batch = std::max<uint64_t>(32, (HOST_BATCH_BYTES / std::max<size_t>(per_shot, 1)) / 32 * 32);
        for (uint64_t s0 = 0; s0 < shots; s0 += batch) {
            const uint64_t n = std::min(batch, shots - s0);
            c->d_in.ensure(n * vs);
            c->d_out.ensure(n * std::max<size_t>(out_row, 1));
            c->d_obs.ensure(n * std::max<size_t>(obs_row, 1));
            if (c->M) {
                if (device) {
                    check(copy_rows_async(c->d_in.p, vs, meas + s0 * meas_stride, meas_stride, c->in_row(), n,
                                           hipMemcpyDeviceToDevice, c->stream), "copy measurements");
                } else {
                    check(copy_rows_async(c->d_in.p, vs, meas + s0 * meas_stride, meas_stride, c->in_row(), n,
                                           hipMemcpyHostToDevice, c->stream), "copy measurements");
                }
            }
            M2dParams P = c->params(n, c->d_in.as<uint8_t>(), vs, c->d_out.as<uint8_t>(), std::max<size_t>(out_row, 1),
                                    n_out, c->d_obs.as<uint8_t>(), std::max<size_t>(obs_row, 1), n_obs);
            if (sweep) {
                c->d_sweep.ensure(n * srow);
                check(copy_rows_async(c->d_sweep.p, srow, sweep + s0 * sweep_stride, sweep_stride, srow, n,
                                       device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream),
                      "copy sweep bits");
                P.sweep = c->d_sweep.as<uint8_t>();
                P.sweep_stride = srow;
            }
            c->launch(P);
            const hipMemcpyKind back = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
            auto deliver = [&](DevBuf &src, size_t row, uint32_t ncols, uint8_t *dst, size_t dst_stride) {
                if (!ncols) return;
                const uint8_t *from = src.as<uint8_t>();
                size_t from_stride = row, width = row;
                if (!packed) {
                    c->d_tmp.ensure(n * ncols);
                    check(launch_unpack_rows(from, row, c->d_tmp.as<uint8_t>(), ncols, n, ncols, 0, c->stream), "unpack");
                    from = c->d_tmp.as<uint8_t>();
                    from_stride = ncols;
                    width = ncols;
                }
                check(copy_rows_async(dst + s0 * dst_stride, dst_stride, from, from_stride, width, n, back, c->stream),
                      "copy results");
                check(hipStreamSynchronize(c->stream), "m2d");
            };
            deliver(c->d_out, std::max<size_t>(out_row, 1), n_out, dets, dets_stride);
            deliver(c->d_obs, std::max<size_t>(obs_row, 1), n_obs, obs, obs_stride);
            check(hipStreamSynchronize(c->stream), "m2d");
        }
    });
}

int mfs_m2d_bench(mfs_m2d *c, uint64_t shots, int warmup, int iters, double out_ms[1]) {
    return guarded([&] {
        // Device-resident packed rows: measurements at the vector stride, detection events with
        // appended observables.
        const size_t vs = std::max<size_t>(c->in_vec_stride(), 16);
        const uint32_t n_out = c->D + c->O;
        const size_t out_row = std::max<size_t>((n_out + 7) / 8, 1);
        DevBuf in, out;
        in.ensure(shots * vs);
        out.ensure(shots * out_row);
        check(hipMemsetAsync(in.p, 0x5A, shots * vs, c->stream), "fill");
        const M2dParams P = c->params(shots, in.as<uint8_t>(), vs, out.as<uint8_t>(), out_row, n_out, nullptr, 0, 0);
        if (!c->ev0) {
            check(hipEventCreate(&c->ev0), "event");
            check(hipEventCreate(&c->ev1), "event");
        }
        for (int i = 0; i < warmup; i++) c->launch(P);
        check(hipStreamSynchronize(c->stream), "warmup");
        double total = 0;
        for (int i = 0; i < iters; i++) {
            check(hipEventRecord(c->ev0, c->stream), "event");
            c->launch(P);
            check(hipEventRecord(c->ev1, c->stream), "event");
            check(hipEventSynchronize(c->ev1), "event");
            float ms = 0;
            check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed");
            total += ms;
        }
        out_ms[0] = iters ? total / iters : 0;
    });
}

}  // extern "C"
