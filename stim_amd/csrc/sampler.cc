// C-ABI entry points of the stim_amd sampler: compile, sample, sample_write, fixed-fault sampling,
// and the result-format writers.
#include <array>
#include <thread>

#include "sampler_impl.h"
#include "tableau.h"

#ifndef STIM_AMD_BUILD_ID
#define STIM_AMD_BUILD_ID "unknown"
#endif

namespace {

enum class Fmt { F01, B8, R8, HITS, DETS, PTB64 };

Fmt parse_format(const char *f) {
    std::string s = f ? f : "01";
    if (s == "01") return Fmt::F01;
    if (s == "b8") return Fmt::B8;
    if (s == "r8") return Fmt::R8;
    if (s == "hits") return Fmt::HITS;
    if (s == "dets") return Fmt::DETS;
    if (s == "ptb64") return Fmt::PTB64;
    throw std::invalid_argument("Unrecognized sample format: '" + s + "'. Expected one of: 01, b8, r8, ptb64, hits, dets");
}

struct FileCloser {
    void operator()(FILE *f) const {
        if (f) fclose(f);
    }
};

// Writes packed shot-major rows (ncols bits each) in a per-shot text or byte format.
// prefix1 applies to columns < transition, prefix2 to the rest (dets format only).
void write_rows(FILE *out, Fmt fmt, const uint8_t *rows, size_t stride, uint64_t nshots, uint32_t ncols, char prefix1,
                char prefix2, uint32_t transition) {
    if (transition == 0) {
        transition = ncols;
        prefix1 = prefix2;
    } else if (prefix1 == prefix2 || transition >= ncols) {
        transition = ncols;
    }
    const size_t row_bytes = (ncols + 7) / 8;
    if (fmt == Fmt::B8 && stride == row_bytes) {
        // Contiguous rows are the file's bytes as they are.
        fwrite(rows, 1, nshots * row_bytes, out);
        return;
    }
    if (fmt == Fmt::F01) {
        // Eight characters per packed byte from a table (bit 0 first), whole chunk per write.
        static const std::array<uint64_t, 256> chars = [] {
            std::array<uint64_t, 256> t{};
            for (uint32_t b = 0; b < 256; b++) {
                for (uint32_t k = 0; k < 8; k++) t[b] |= (uint64_t)('0' + ((b >> k) & 1)) << (8 * k);
            }
            return t;
        }();
        const size_t line = (size_t)ncols + 1;
        const uint64_t per = std::max<uint64_t>(1, ((size_t)16 << 20) / line);
        std::vector<char> text;
        for (uint64_t s0 = 0; s0 < nshots; s0 += per) {
            const uint64_t n = std::min(per, nshots - s0);
            text.resize(n * line + 8);
            for (uint64_t s = 0; s < n; s++) {
                const uint8_t *r = rows + (s0 + s) * stride;
                char *o = text.data() + s * line;
                const uint32_t full = ncols / 8;
                for (uint32_t b = 0; b < full; b++) std::memcpy(o + 8 * b, &chars[r[b]], 8);
                for (uint32_t m = 8 * full; m < ncols; m++) o[m] = (char)('0' + ((r[m >> 3] >> (m & 7)) & 1));
                o[ncols] = '\n';
            }
            fwrite(text.data(), 1, n * line, out);
        }
        return;
    }
    for (uint64_t s = 0; s < nshots; s++) {
        const uint8_t *r = rows + s * stride;
        auto bit = [&](uint32_t m) { return (r[m >> 3] >> (m & 7)) & 1; };
        switch (fmt) {
            case Fmt::B8:
                fwrite(r, 1, row_bytes, out);
                break;
            case Fmt::F01:
                break;
            case Fmt::HITS: {
                bool first = true;
                for (uint32_t m = 0; m < ncols; m++) {
                    if ((m & 7) == 0 && m + 8 <= ncols && r[m >> 3] == 0) {
                        m += 7;
                        continue;
                    }
                    if (bit(m)) {
                        if (!first) putc(',', out);
                        first = false;
                        fprintf(out, "%llu", (unsigned long long)m);
                    }
                }
                putc('\n', out);
                break;
            }
            case Fmt::DETS: {
                fputs("shot", out);
                for (uint32_t m = 0; m < ncols; m++) {
                    if ((m & 7) == 0 && m + 8 <= ncols && r[m >> 3] == 0) {
                        m += 7;
                        continue;
                    }
                    if (bit(m)) {
                        if (m < transition) fprintf(out, " %c%u", prefix1, m);
                        else fprintf(out, " %c%u", prefix2, m - transition);
                    }
                }
                putc('\n', out);
                break;
            }
            case Fmt::R8: {
                uint32_t run = 0;
                for (uint32_t m = 0; m < ncols; m++) {
                    if (bit(m)) {
                        putc((int)run, out);
                        run = 0;
                    } else {
                        run++;
                        if (run == 255) {
                            putc(255, out);
                            run = 0;
                        }
                    }
                }
                putc((int)run, out);
                break;
            }
            case Fmt::PTB64:
                break;
        }
    }
}

// ptb64: for each group of 64 shots, for each column, 8 bytes holding that column's bit for the 64 shots.
void write_ptb64(FILE *out, const uint8_t *rows, size_t stride, uint64_t nshots, uint32_t ncols) {
    for (uint64_t g = 0; g < nshots / 64; g++) {
        for (uint32_t m = 0; m < ncols; m++) {
            uint64_t v = 0;
            for (uint32_t k = 0; k < 64; k++) {
                const uint8_t *r = rows + (g * 64 + k) * stride;
                v |= (uint64_t)((r[m >> 3] >> (m & 7)) & 1) << k;
            }
            uint8_t b[8];
            for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
            fwrite(b, 1, 8, out);
        }
    }
}

thread_local std::string g_last_error;

}  // namespace

int mfs_detail::record_error(const std::exception &e, int kind) {
    g_last_error = e.what();
    return kind;
}

extern "C" {

const char *mfs_last_error(void) {
    return g_last_error.c_str();
}

int mfs_compile(const char *stim_text, size_t len, uint64_t seed, int has_seed, mfs_sampler **out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = nullptr;
        auto s = std::make_unique<mfs_sampler>();
        s->text.assign(stim_text ? stim_text : "", stim_text ? len : 0);
        s->block = parse_circuit(s->text.data(), s->text.size());
        s->flat = flatten(s->block);
        s->prog = compile_program(s->flat);
        if (has_seed) {
            s->key = splitmix64(seed ^ 0xDEADBEEF124Cull);
        } else {
            std::random_device rd;
            s->key = ((uint64_t)rd() << 32) ^ rd();
        }
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            throw HipError("no HIP device available: the stim_amd sampler runs only on the GPU");
        }
        check(hipStreamCreate(&s->stream), "stream");
        s->pick_tile_words(0);
        s->upload_program();
        s->probe_randomization();
        *out = s.release();
    });
}

int mfs_compile_measurements(const char *stim_text, size_t len, uint64_t seed, int has_seed, int skip_reference_sample,
                             const uint8_t *reference_sample_bits, size_t reference_len, mfs_sampler **out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = nullptr;
        if (skip_reference_sample && reference_sample_bits)
            throw std::invalid_argument("skip_reference_sample = True but reference_sample is not None.");
        auto s = std::make_unique<mfs_sampler>();
        s->text.assign(stim_text ? stim_text : "", stim_text ? len : 0);
        s->block = parse_circuit(s->text.data(), s->text.size());
        s->flat = flatten(s->block);
        s->prog = compile_program(s->flat, true);
        const uint64_t M = s->prog.num_measurements;
        std::vector<uint32_t> words((M + 31) / 32, 0);
        if (reference_sample_bits) {
            if (reference_len < (M + 7) / 8) throw std::invalid_argument("reference_sample is shorter than the circuit's measurements");
            for (uint64_t m = 0; m < M; m++) {
                if ((reference_sample_bits[m >> 3] >> (m & 7)) & 1) words[m >> 5] |= 1u << (m & 31);
            }
        } else if (!skip_reference_sample) {
            const std::vector<uint8_t> ref = reference_sample(s->flat);
            for (uint64_t m = 0; m < M; m++) {
                if (ref[m]) words[m >> 5] |= 1u << (m & 31);
            }
        }
        if (has_seed) {
            s->key = splitmix64(seed ^ 0xDEADBEEF124Cull);
        } else {
            std::random_device rd;
            s->key = ((uint64_t)rd() << 32) ^ rd();
        }
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            throw HipError("no HIP device available: the stim_amd sampler runs only on the GPU");
        }
        check(hipStreamCreate(&s->stream), "stream");
        s->meas_mode = true;
        s->engine_mode = 0;
        s->invert_words = words;
        s->pick_tile_words(0);
        s->upload_program();
        s->probe_randomization();
        bool any = false;
        for (uint32_t w : words) any |= w != 0;
        if (any) {
            mfs_sampler::upload(s->d_invert, words, "reference sample");
            s->has_invert = true;
        }
        *out = s.release();
    });
}

int mfs_dem_compile(const char *dem_text, size_t len, uint64_t seed, int has_seed, mfs_sampler **out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = nullptr;
        const DemModel model = parse_dem(dem_text, dem_text ? len : 0);
        auto s = std::make_unique<mfs_sampler>();
        if (has_seed) {
            s->key = splitmix64(seed ^ 0xDEADBEEF124Cull);
        } else {
            std::random_device rd;
            s->key = ((uint64_t)rd() << 32) ^ rd();
        }
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            throw HipError("no HIP device available: the stim_amd sampler runs only on the GPU");
        }
        check(hipStreamCreate(&s->stream), "stream");
        s->init_dem(model);
        *out = s.release();
    });
}

int mfs_dem_sample_errors(mfs_sampler *s, uint64_t shots, uint32_t flags, uint8_t *dets, size_t dets_stride,
                          uint8_t *obs, size_t obs_stride, uint8_t *errs, size_t errs_stride) {
    return guarded([&] {
        if (!s->dem_mode) throw std::invalid_argument("only a detector error model sampler returns errors");
        if (flags & (MFS_PREPEND_OBSERVABLES | MFS_APPEND_OBSERVABLES))
            throw std::invalid_argument("errors are returned with separate observables only");
        const uint64_t E = s->dem_col.size();
        const size_t erow = (flags & MFS_BIT_PACKED) ? (E + 7) / 8 : E;
        if (E && (!errs || errs_stride < erow)) throw std::invalid_argument("error rows are too short");
        s->sample(shots, flags | MFS_SEPARATE_OBSERVABLES, dets, dets_stride, obs, obs_stride, nullptr, nullptr, 0, false,
                  E ? errs : nullptr, errs_stride);
    });
}

int mfs_dem_parse_stats(const char *dem_text, size_t len, uint64_t out[3]) {
    return guarded([&] {
        const DemModel model = parse_dem(dem_text, dem_text ? len : 0);
        out[0] = model.num_detectors;
        out[1] = model.num_observables;
        out[2] = model.errors.size();
    });
}

void mfs_destroy(mfs_sampler *s) {
    delete s;
}

int mfs_parse_stats(const char *stim_text, size_t len, uint64_t out[10]) {
    return guarded([&] {
        std::string text(stim_text ? stim_text : "", stim_text ? len : 0);
        Block b = parse_circuit(text.data(), text.size());
        std::vector<const Instr *> flat = flatten(b);
        Program prog = compile_program(flat);
        out[0] = prog.stats.num_qubits;
        out[1] = prog.stats.num_measurements;
        out[2] = prog.stats.num_detectors;
        out[3] = prog.stats.num_observables;
        out[4] = prog.stats.max_lookback;
        out[5] = prog.ops.size();
        out[6] = prog.num_syncs;
        out[7] = prog.num_noise_ops;
        out[8] = prog.ring;
        out[9] = lds_bytes(prog, 2);
    });
}

int mfs_expand_plan_stats(const char *stim_text, size_t len, uint64_t out[8]) {
    return guarded([&] {
        std::string text(stim_text ? stim_text : "", stim_text ? len : 0);
        Block b = parse_circuit(text.data(), text.size());
        std::vector<const Instr *> flat = flatten(b);
        const Stats st = compute_stats(flat);
        RefStats rs;
        (void)reference_sample(flat, &rs);
        const ExpandPlan E = plan_expansion(flat, (uint32_t)st.num_measurements, (uint32_t)st.num_detectors,
                                            (uint32_t)st.num_observables, rs.random_recorded, rs.random_total);
        out[0] = E.complete ? 1 : 0;
        out[1] = st.num_measurements;
        out[2] = E.pivots;
        out[3] = E.complete ? E.free_slots.size() : (uint64_t)st.num_measurements - E.pivots;
        out[4] = rs.random_recorded;
        out[5] = rs.random_total;
        out[6] = E.phases;
        out[7] = E.levels;
    });
}

int mfs_expand_shots_host(const char *stim_text, size_t len, uint64_t shots, const uint8_t *y, size_t y_stride,
                          const uint8_t *free_bits, size_t free_stride, uint8_t *out, size_t out_stride) {
    return guarded([&] {
        std::string text(stim_text ? stim_text : "", stim_text ? len : 0);
        Block b = parse_circuit(text.data(), text.size());
        std::vector<const Instr *> flat = flatten(b);
        const Stats st = compute_stats(flat);
        RefStats rs;
        (void)reference_sample(flat, &rs);
        const ExpandPlan E = plan_expansion(flat, (uint32_t)st.num_measurements, (uint32_t)st.num_detectors,
                                            (uint32_t)st.num_observables, rs.random_recorded, rs.random_total);
        if (!E.complete) throw std::invalid_argument("expansion unavailable: " + E.why);
        for (uint64_t s = 0; s < shots; s++)
            expand_shot_host(E, y + s * y_stride, free_bits + s * free_stride, out + s * out_stride);
    });
}

int mfs_reference_sample(const char *stim_text, size_t len, uint8_t *out, size_t out_len) {
    return guarded([&] {
        std::string text(stim_text ? stim_text : "", stim_text ? len : 0);
        Block b = parse_circuit(text.data(), text.size());
        const std::vector<uint8_t> ref = reference_sample(flatten(b));
        const size_t need = (ref.size() + 7) / 8;
        if (need && !out) throw std::invalid_argument("out is null");
        if (out_len < need) throw std::invalid_argument("output buffer too small for the reference sample");
        std::memset(out, 0, need);
        for (size_t m = 0; m < ref.size(); m++) {
            if (ref[m]) out[m >> 3] |= (uint8_t)(1u << (m & 7));
        }
    });
}

int mfs_circuit_stats(const mfs_sampler *s, uint64_t out[5]) {
    return guarded([&] {
        const Stats &st = s->prog.stats;
        out[0] = st.num_qubits;
        out[1] = st.num_measurements;
        out[2] = st.num_detectors;
        out[3] = st.num_observables;
        out[4] = st.max_lookback;
    });
}

int mfs_sample(mfs_sampler *s, uint64_t shots, uint32_t flags, uint8_t *dets, size_t dets_stride, uint8_t *obs,
               size_t obs_stride) {
    return guarded([&] { s->sample(shots, flags, dets, dets_stride, obs, obs_stride); });
}

int mfs_sample_with_faults(mfs_sampler *s, uint64_t shots, const uint64_t *shot_offsets, const uint32_t *faults,
                           uint32_t flags, uint8_t *dets, size_t dets_stride, uint8_t *obs, size_t obs_stride) {
    return guarded([&] {
        if (!shot_offsets) throw std::invalid_argument("shot_offsets is null");
        static const uint32_t none[3] = {0, 0, 0};
        s->sample(shots, flags, dets, dets_stride, obs, obs_stride, shot_offsets, faults ? faults : none);
    });
}

int mfs_sample_write(mfs_sampler *s, uint64_t shots, const char *path, const char *format, int prepend_observables,
                     int append_observables, const char *obs_path, const char *obs_format) {
    return guarded([&] {
        const Fmt fmt = parse_format(format);
        const bool has_obs = obs_path && obs_path[0];
        const Fmt ofmt = has_obs ? parse_format(obs_format) : Fmt::F01;
        if ((prepend_observables ? 1 : 0) + (append_observables ? 1 : 0) + (has_obs ? 1 : 0) > 1) {
            throw std::out_of_range("Can't combine --prepend_observables, --append_observables, or --obs_out");
        }
        if ((fmt == Fmt::PTB64 || (has_obs && ofmt == Fmt::PTB64)) && shots % 64) {
            throw std::invalid_argument("shots must be a multiple of 64 to use ptb64 format.");
        }
        std::unique_ptr<FILE, FileCloser> out(fopen(path, "wb"));
        if (!out) throw std::invalid_argument(std::string("Failed to open '") + path + "' to write.");
        std::unique_ptr<FILE, FileCloser> oout;
        if (has_obs) {
            oout.reset(fopen(obs_path, "wb"));
            if (!oout) throw std::invalid_argument(std::string("Failed to open '") + obs_path + "' to write.");
        }
        const uint32_t D = s->prog.num_detectors, O = s->prog.num_observables;
        uint32_t flags = MFS_BIT_PACKED | (prepend_observables ? MFS_PREPEND_OBSERVABLES : 0) |
                         (append_observables ? MFS_APPEND_OBSERVABLES : 0) | (has_obs ? MFS_SEPARATE_OBSERVABLES : 0);
        const uint32_t ncols = D + O * ((prepend_observables ? 1 : 0) + (append_observables ? 1 : 0));
        const size_t drow = (ncols + 7) / 8, orow = (O + 7) / 8;
        const uint64_t chunk = std::max<uint64_t>(64, ((64ull << 20) / std::max<size_t>(drow + orow, 1)) / 64 * 64);
        char p1 = s->meas_mode ? 'M' : 'D', p2 = 'L';
        uint32_t transition = D;
        if (prepend_observables) {
            p1 = 'L';
            p2 = 'D';
            transition = O;
        }
        // A writer thread formats and writes chunk c while the GPU samples chunk c + 1 into the other
        // buffer pair.
        std::vector<uint8_t> dbufs[2], obufs[2];
        std::thread writer;
        struct Join {
            std::thread &t;
            ~Join() {
                if (t.joinable()) t.join();
            }
        } join_writer{writer};
        const size_t dstride = std::max<size_t>(drow, 1), ostride = std::max<size_t>(orow, 1);
        uint32_t c = 0;
        for (uint64_t s0 = 0; s0 < shots; s0 += chunk, c ^= 1) {
            const uint64_t n = std::min(chunk, shots - s0);
            std::vector<uint8_t> &db = dbufs[c], &ob = obufs[c];
            db.assign(n * dstride, 0);
            ob.assign(n * ostride, 0);
            s->sample(n, flags, db.data(), dstride, ob.data(), ostride);
            if (writer.joinable()) writer.join();
            writer = std::thread([&, n, c] {
                FILE *f = out.get();
                const std::vector<uint8_t> &d = dbufs[c], &o = obufs[c];
                if (fmt == Fmt::PTB64) write_ptb64(f, d.data(), dstride, n, ncols);
                else write_rows(f, fmt, d.data(), dstride, n, ncols, p1, p2, transition);
                if (has_obs) {
                    if (ofmt == Fmt::PTB64) write_ptb64(oout.get(), o.data(), ostride, n, O);
                    else write_rows(oout.get(), ofmt, o.data(), ostride, n, O, 'L', 'L', O);
                }
            });
        }
    });
}

int mfs_bench_device(mfs_sampler *s, uint64_t shots, int warmup, int iters, double out_ms[2]) {
    return guarded([&] {
        mfs_sampler::Layout L = s->layout(MFS_BIT_PACKED | MFS_SEPARATE_OBSERVABLES);
        DevBuf dd, od;
        dd.ensure(shots * std::max<size_t>(L.det_row, 1));
        od.ensure(shots * std::max<size_t>(L.obs_row, 1));
        if (!s->ev0) {
            check(hipEventCreate(&s->ev0), "event");
            check(hipEventCreate(&s->ev1), "event");
        }
        const uint64_t tiles = (shots + s->W() - 1) / s->W();
        const bool expand = s->meas_mode && s->use_expansion(false, shots);
        const bool sym = !expand && s->use_symptoms(false, shots);
        // Expansion: detection events into padded y rows, then the expansion kernel (timed alone in
        // out_ms[1] with an event between the two).
        DevBuf yb;
        hipEvent_t evm = nullptr;
        struct EvGuard {
            hipEvent_t &e;
            ~EvGuard() {
                if (e) (void)hipEventDestroy(e);
            }
        } evm_guard{evm};
        if (expand) {
            yb.ensure(shots * s->det_src->row_pitch_override);
            check(hipEventCreate(&evm), "event");
        }
        double expand_total = 0;
        auto one = [&]() {
            if (expand) {
                mfs_sampler &d = *s->det_src;
                const mfs_sampler::Layout DL = d.layout(MFS_BIT_PACKED | MFS_APPEND_OBSERVABLES);
                d.tiles_used += d.sym_launch(DL, shots, yb.as<uint8_t>(), d.row_pitch_override, nullptr, 0, d.tiles_used,
                                             nullptr, nullptr);
                check(hipEventRecord(evm, s->stream), "event");
                check(launch_expand(s->expand_params(shots, yb.as<uint8_t>(), dd.as<uint8_t>(), L.det_row, s->expand_tiles),
                                    s->stream),
                      "expansion launch");
                s->expand_tiles += (shots + 31) / 32;
                return;
            }
            if (sym) {
                s->tiles_used += s->sym_launch(L, shots, dd.as<uint8_t>(), L.det_row, od.as<uint8_t>(), L.obs_row,
                                               s->tiles_used, nullptr, nullptr);
                return;
            }
            LaunchParams P = s->base_params();
            s->outputs(L, dd.as<uint8_t>(), L.det_row, od.as<uint8_t>(), L.obs_row, P);
            P.shots = shots;
            P.tile0 = s->tiles_used;
            P.mode = MODE_SAMPLED;
            P.randomize = s->randomize ? 1 : 0;
            s->launch(P, tiles);
            s->tiles_used += tiles;
        };
        for (int i = 0; i < warmup; i++) one();
        check(hipStreamSynchronize(s->stream), "warmup");
        double kernel_total = 0;
        for (int i = 0; i < iters; i++) {
            check(hipEventRecord(s->ev0, s->stream), "event");
            one();
            check(hipEventRecord(s->ev1, s->stream), "event");
            check(hipEventSynchronize(s->ev1), "event");
            float ms = 0;
            check(hipEventElapsedTime(&ms, s->ev0, s->ev1), "elapsed");
            kernel_total += ms;
            if (expand) {
                float xm = 0;
                check(hipEventElapsedTime(&xm, evm, s->ev1), "elapsed");
                expand_total += xm;
            }
        }
        out_ms[0] = iters ? kernel_total / iters : 0;
        out_ms[1] = expand ? (iters ? expand_total / iters : 0) : out_ms[0];
    });
}

int mfs_set_tile_words(mfs_sampler *s, int words) {
    int rc = guarded([&] {
        if (s->dem_mode) throw std::invalid_argument("a detector error model sampler has no frame engine");
        if (words != 0 && words != 1 && words != 2 && words != 4) throw std::invalid_argument("tile words must be 0, 1, 2 or 4");
        s->pick_tile_words(words);
    });
    if (rc) return -rc;
    return s->T;
}

int mfs_set_engine(mfs_sampler *s, int engine) {
    return guarded([&] {
        if (engine < 0 || engine > 2) throw std::invalid_argument("engine must be 0 (auto), 1 (frame) or 2 (symptom)");
        if (s->dem_mode) {
            if (engine == 1) throw std::invalid_argument("a detector error model sampler has no frame engine");
            return;
        }
        if (s->meas_mode) {
            // 2 = expansion of detection events (the measurement sampler's symptom-engine path).
            if (engine == 2 && !s->ensure_expansion())
                throw std::invalid_argument("expansion unavailable: " + s->expand_why);
            s->engine_mode = engine;
            return;
        }
        s->engine_mode = engine;
        if (engine == 2 && !s->use_symptoms(false)) throw std::invalid_argument("symptom engine unavailable");
    });
}

int mfs_engine_info(mfs_sampler *s, uint64_t out[10]) {
    return guarded([&] {
        if (s->meas_mode && s->use_expansion(false, mfs_sampler::ANY_SHOTS)) {
            // Expansion: the symptom-engine figures are the detector sampler's.
            mfs_sampler *d = s->det_src.get();
            out[0] = 3;
            out[1] = d->plan.classes.size();
            out[2] = d->plan.columns.size();
            out[3] = d->sym_entries;
            out[4] = (uint64_t)(d->sym_build_ms * 1000.0);
            out[5] = d->plan.total_sites;
            out[6] = s->expand_free;
            out[7] = s->expand_phases;
            out[8] = s->expand_items;
            out[9] = 0;
            return;
        }
        const bool sym = s->use_symptoms(false);
        out[0] = sym ? 2 : 1;
        out[1] = s->plan.classes.size();
        out[2] = s->plan.columns.size();
        out[3] = s->sym_entries;
        out[4] = (uint64_t)(s->sym_build_ms * 1000.0);
        out[5] = s->plan.total_sites;
        out[6] = out[7] = out[8] = out[9] = 0;
        if (sym && s->sym_fast) {
            out[8] = s->orec_blocks;
            out[9] = s->orec_records;
        }
        if (sym) {
            // Tile of a bit-packed sample with separate observables.
            static uint8_t dummy;
            const auto L = s->layout(MFS_BIT_PACKED | MFS_SEPARATE_OBSERVABLES | MFS_DEVICE_OUTPUT);
            const auto r = s->sym_prepare(L, 1, &dummy, L.det_row, &dummy, L.obs_row, 0);
            out[6] = r.P.shots_per_tile;
            out[7] = r.lds_scan;
        }
    });
}

int mfs_set_symptom_tile_kb(mfs_sampler *s, int kb) {
    return guarded([&] {
        if (kb == 0 || kb < -128 || kb > 128) throw std::invalid_argument("symptom tile size must be 1..128 KB");
        // A negative size selects the per-lane-walk queue kernel instead of the scan kernel.
        if (kb < 0 && s->dem_mode)
            throw std::invalid_argument("a detector error model sampler samples only with the scan kernel");
        s->sym_scan = kb > 0;
        s->sym_tile_kb = kb > 0 ? kb : -kb;
    });
}

int mfs_set_scan_variant(mfs_sampler *s, int variant) {
    return guarded([&] {
        if (variant != 0 && variant != 1) throw std::invalid_argument("scan variant must be 0 (automatic) or 1 (generic)");
        s->sym_variant = variant;
    });
}

int mfs_set_stream_position(mfs_sampler *s, uint64_t tiles) {
    return guarded([&] { s->tiles_used = tiles; });
}

int mfs_get_stream_position(const mfs_sampler *s, uint64_t *tiles) {
    return guarded([&] { *tiles = s->tiles_used; });
}

int mfs_rng_counter(int kernel, uint64_t key, uint64_t tile, uint32_t lane, uint32_t draw, uint32_t tag,
                    uint32_t out[6]) {
    return guarded([&] {
        if (kernel != 0 && kernel != 1) throw std::invalid_argument("kernel must be 0 (symptom) or 1 (frame)");
        if (kernel == 0 && lane > 255) throw std::invalid_argument("symptom kernels have 256 lanes");
        out[0] = draw;
        out[1] = kernel == 0 ? sym_lane_word(lane, tile) : lane;
        out[2] = (uint32_t)tile;
        out[3] = tag;
        out[4] = (uint32_t)key;
        out[5] = kernel == 0 ? (uint32_t)(key >> 32) : frame_key_hi((uint32_t)(key >> 32), tile);
    });
}

int mfs_philox(const uint32_t *counters, uint64_t n, uint64_t key, uint32_t *out) {
    return guarded([&] {
        if (!n) return;
        if (n > (1ull << 24)) throw std::invalid_argument("at most 2^24 counters per call");
        DevBuf d_in, d_out;
        d_in.ensure(16 * n);
        d_out.ensure(16 * n);
        check(hipMemcpy(d_in.p, counters, 16 * n, hipMemcpyHostToDevice), "hipMemcpy");
        check(launch_philox(d_in.as<uint32_t>(), n, key, d_out.as<uint32_t>(), nullptr), "philox kernel");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check(hipMemcpy(out, d_out.p, 16 * n, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

const char *mfs_build_id(void) {
    return STIM_AMD_BUILD_ID;
}

int mfs_get_info(const mfs_sampler *s, uint64_t out[8]) {
    return guarded([&] {
        out[0] = (uint64_t)s->T;
        out[1] = s->lds;
        out[2] = s->prog.ops.size();
        out[3] = s->prog.num_syncs;
        out[4] = s->prog.num_noise_ops;
        out[5] = s->randomize ? 1 : 0;
        out[6] = s->prog.ring;
        out[7] = s->prog.num_qubits;
    });
}

}  // extern "C"

