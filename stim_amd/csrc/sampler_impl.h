// The sampler object behind the C ABI (shared by sampler.cc, batch.cc and server.cc).
//
// Mirrors the reference's CompiledDetectorSampler (src/stim/py/compiled_detector_sampler.pybind.{h,cc})
// and its drivers (src/stim/simulators/frame_simulator_util.inl): the circuit is compiled once into a
// device program; each call simulates batches of shot tiles on the GPU, writing shot-major rows
// directly (no separate transpose pass), and copies them to the caller.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/stim_amd.h"
#include "circuit.h"
#include "compiler.h"
#include "program.h"
#include "dem.h"
#include "symptom.h"
#include "rng.h"
#include "host_pipe.h"
#include "expand.h"
#include "m2d.h"
#include "tableau.h"

namespace sa {
hipError_t launch_frame_kernel(int T, const DevProgram &prog, const LaunchParams &P, uint32_t num_tiles, size_t lds,
                               hipStream_t stream);
hipError_t launch_pack_rows(const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride, uint64_t rows,
                            uint32_t ncols, hipStream_t stream);
hipError_t launch_sym_sample(const SymDevParams &P, uint32_t tiles, size_t lds, hipStream_t stream);
hipError_t launch_sym_fixed(const SymDevParams &P, const uint64_t *off, const uint32_t *cols, uint32_t tiles,
                            size_t lds, hipStream_t stream);
hipError_t launch_sym_scan(const SymDevParams &P, uint32_t tiles, size_t lds, hipStream_t stream, bool fast);
hipError_t launch_sym_batch(SymDevParams *jobs, const uint64_t *prefix, const uint64_t *advance, uint32_t njobs,
                            uint32_t *counter, uint32_t grid, size_t lds, hipStream_t stream);
int sym_batch_blocks_per_cu(size_t lds);
hipError_t launch_sym_count(const uint8_t *det, uint32_t det_row, const uint8_t *obs, uint32_t obs_row, uint64_t rows,
                            uint32_t *counts, hipStream_t stream);
hipError_t launch_sym_fill(const uint8_t *det, uint32_t det_row, const uint8_t *obs, uint32_t obs_row, uint64_t rows,
                           uint32_t D, const uint32_t *offs, uint32_t *ent, hipStream_t stream);
hipError_t launch_unpack_rows(const uint8_t *src, uint64_t src_stride, uint8_t *dst, uint64_t dst_stride, uint64_t rows,
                              uint32_t ncols, uint32_t base_char, hipStream_t stream);
hipError_t launch_philox(const uint32_t *ctr, uint64_t n, uint64_t key, uint32_t *out, hipStream_t stream);
hipError_t launch_expand(const ExpandDev &E, hipStream_t stream);
size_t m2d_lds_bytes(uint32_t blocks, uint32_t sweep_blocks);
}  // namespace sa

using namespace sa;

namespace mfs_detail {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        throw HipError(std::string(what) + ": " + hipGetErrorString(e));
    }
}

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void ensure(size_t n) {
        if (n <= cap) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        check(hipMalloc(&p, std::max<size_t>(n, 256)), "hipMalloc");
        cap = std::max<size_t>(n, 256);
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <class X>
    X *as() {
        return reinterpret_cast<X *>(p);
    }
};

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

constexpr size_t MAX_LDS = LDS_PER_WORKGROUP;
constexpr size_t HOST_BATCH_BYTES = (size_t)1 << 30;
// Threads of an expansion workgroup (m2d_kernel.hip's workgroup size), the width of its schedule.
constexpr uint32_t EXPAND_THREADS = 1024;
// Host-bound bytes per batch of a call that returns rows in host memory (one of two pipeline slots).
constexpr size_t PIPE_BATCH_BYTES = (size_t)64 << 20;

// Records the message of a failed C-ABI call (read back by mfs_last_error) and returns its kind.
int record_error(const std::exception &e, int kind);

// Runs f, mapping exceptions to the ABI's error kinds (std::invalid_argument -> ValueError,
// std::out_of_range -> IndexError, anything else -> RuntimeError).
template <class F>
int guarded(F f) {
    try {
        f();
        return MFS_OK;
    } catch (const std::out_of_range &e) {
        return record_error(e, MFS_ERR_OUT_OF_RANGE);
    } catch (const std::invalid_argument &e) {
        return record_error(e, MFS_ERR_INVALID_ARGUMENT);
    } catch (const std::exception &e) {
        return record_error(e, MFS_ERR_RUNTIME);
    }
}

}  // namespace mfs_detail

using namespace mfs_detail;

struct mfs_sampler {
    std::string text;
    Block block;
    std::vector<const Instr *> flat;
    Program prog;
    uint64_t key = 0;
    uint64_t tiles_used = 0;
    int T = 2;
    size_t lds = 0;
    bool randomize = false;
    hipStream_t stream = nullptr;
    DevBuf d_blob, d_ntgt, d_otab, d_faults, d_fault_off, d_out_det, d_out_obs, d_tmp;
    // Symptom engine state (built on first use).
    int engine_mode = 0;         // 0 automatic, 1 frame engine, 2 symptom engine
    int sym_state = 0;           // 0 not built, 1 ready, -1 unsupported
    std::string sym_why;
    SymPlan plan;
    uint64_t sym_entries = 0;
    double sym_build_ms = 0;
    DevBuf d_classes, d_sotab, d_recs, d_ovf, d_sym_tmp, d_fix_off, d_fix_cols, d_recs32, d_ovf32, d_groups, d_thr;
    // Detector-error-model samplers: sampled site -> mechanism index, and the fired-mechanism rows of
    // the launch in progress (set only while a call returns errors).
    DevBuf d_site_mech, d_err, d_err_tmp;
    // Outcome records of the specialised scan kernel (built with the symptom table when it applies).
    DevBuf d_orecs, d_orec_base, d_orec_sites, d_orec_ovf;
    uint64_t orec_blocks = 0, orec_records = 0;
    bool sym_fast = false;
    int sym_variant = 0;         // 0: specialised scan kernel where it applies; 1: generic scan kernel
    uint32_t *err_words_cur = nullptr;
    uint64_t err_stride_cur = 0;
    // Measurement samplers: the outputs are the measurement results (reference sample bits XORed
    // into the frame's flips through det_invert). They sample on the frame engine, or, where the
    // circuit's detectors and observables fix every deterministic parity, by expanding detection events
    // drawn on the symptom engine into records (plan and semantics in expand.h).
    bool meas_mode = false;
    DevBuf d_invert;
    bool has_invert = false;
    std::vector<uint32_t> invert_words;   // host copy of the inversion (reference sample) words
    int expand_state = 0;                 // 0 not planned, 1 ready, -1 unavailable
    std::string expand_why;
    std::unique_ptr<mfs_sampler> det_src; // detector sampler of the same circuit, on this sampler's stream
    bool borrowed_stream = false;         // the stream belongs to the sampler that owns this one
    size_t row_pitch_override = 0;        // symptom-engine rows padded to this pitch (0: packed width)
    uint64_t expand_key = 0, expand_tiles = 0;
    uint32_t expand_free = 0, expand_phases = 0, expand_items = 0;
    DevBuf d_x_prim, d_x_extra_off, d_x_extra, d_x_free, d_x_sched, d_x_items, d_x_terms, d_x_inv, d_x_y;
    // Detector-error-model samplers: no circuit; the symptom tables come from the model's mechanisms.
    bool dem_mode = false;
    std::vector<uint32_t> dem_col;   // mechanism index (model order) -> symptom column
    bool sym_compact = false;
    bool sym_ovf32 = false;      // some compact record overflows
    bool plan_counted = false;
    uint64_t column_count = 0;
    int sym_tile_kb = 32;        // target LDS image size per symptom tile
    int sym_scan = 1;            // scan kernel (0: per-lane-walk queue kernel)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Device staging of a call comes from the sampler's own buffers (grown with hipMalloc as needed),
    // or, while `arena` is set (the allocation-free entry point), from consecutive 256-byte-aligned
    // pieces of a fixed device region, reused from its start by every batch of shots.
    struct Arena {
        uint8_t *base = nullptr;
        size_t cap = 0, used = 0;
    };
    Arena *arena = nullptr;
    // Calls that return rows in host memory: double-buffered device slots and pinned buffers.
    HostPipe pipe;
    // Points staging at the pipeline's next device slot for one batch (restores the previous arena and,
    // when the call unwinds, ends what the pipeline has in flight).
    struct SlotScope {
        mfs_sampler *s;
        Arena slot{};
        Arena *prev;
        explicit SlotScope(mfs_sampler *o) : s(o), prev(o->arena) {
            slot.base = s->pipe.next_slot(&slot.cap);
            s->arena = &slot;
        }
        ~SlotScope() { s->arena = prev; }
    };
    struct PipeCall {
        mfs_sampler *s;
        bool on;
        ~PipeCall() {
            if (on) s->pipe.drain_noexcept();
        }
    };

    uint8_t *staging(DevBuf &b, size_t n) {
        if (!arena) {
            b.ensure(n);
            return b.as<uint8_t>();
        }
        const size_t at = (arena->used + 255) & ~(size_t)255;
        if (at + n > arena->cap) throw std::invalid_argument("staging exceeds the fixed scratch region");
        arena->used = at + n;
        return arena->base + at;
    }
    void staging_reset() {
        if (arena) arena->used = 0;
    }
    // Shots per batch when every batch stages `row_bytes` per shot in at most three pieces: the
    // arena's capacity, in whole units of `unit` shots (the RNG tile), so splitting a call into batches
    // never changes its tile numbering. Without an arena, `limit`.
    uint64_t arena_batch(uint64_t limit, size_t row_bytes, uint64_t unit) const {
        if (!arena || row_bytes == 0) return limit;
        const size_t usable = arena->cap > 3 * 256 ? arena->cap - 3 * 256 : 0;
        const uint64_t fit = usable / row_bytes / unit * unit;
        if (fit == 0)
            throw std::invalid_argument("rows too wide for the fixed scratch region (" + std::to_string(row_bytes) +
                                        " bytes of staging per shot, " + std::to_string(unit) + " shots per tile)");
        return std::min(limit, fit);
    }

    ~mfs_sampler() {
        if (det_src) det_src.reset();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream && !borrowed_stream) (void)hipStreamDestroy(stream);
    }

    uint32_t W() const { return 32u * (uint32_t)T; }

    void pick_tile_words(int want) {
        if (want) {
            size_t need = lds_bytes(prog, want);
            if (need > MAX_LDS) throw std::invalid_argument("tile does not fit in LDS for this circuit");
            T = want;
        } else if (lds_bytes(prog, 2) <= 64 * 1024) {
            T = 2;
        } else if (lds_bytes(prog, 1) <= MAX_LDS) {
            T = 1;
        } else {
            throw std::invalid_argument(
                "circuit's frame does not fit in LDS (" + std::to_string(prog.num_qubits) + " qubits, record ring " +
                std::to_string(prog.ring) + " rows)");
        }
        lds = lds_bytes(prog, T);
    }

    LaunchParams base_params() {
        LaunchParams P{};
        P.seed = key;
        P.blob0_len = prog.blob0_len;
        P.num_qubits = prog.num_qubits;
        P.ring = prog.ring;
        P.num_obs = prog.num_observables;
        P.det_invert = has_invert ? d_invert.as<uint32_t>() : nullptr;
        return P;
    }

    void launch(const LaunchParams &P, uint64_t tiles) {
        if (tiles >= (1ull << 31)) throw std::invalid_argument("too many shots in one launch");
        DevProgram dp{d_blob.as<uint32_t>(), d_ntgt.as<uint32_t>(), d_otab.as<uint32_t>()};
        check(launch_frame_kernel(T, dp, P, (uint32_t)tiles, lds, stream), "frame kernel launch");
    }

    // Output layout of one sample call.
    struct Layout {
        uint32_t ncols;      // entries per dets row
        size_t det_row;      // bytes per dets row
        size_t obs_row;      // bytes per obs row (separate)
        bool packed;
        bool prepend, append, separate;
        bool via_unpacked;   // packed with a non-byte-aligned dets/obs boundary
    };

    Layout layout(uint32_t flags) const {
        Layout L{};
        L.packed = flags & MFS_BIT_PACKED;
        L.prepend = flags & MFS_PREPEND_OBSERVABLES;
        L.append = flags & MFS_APPEND_OBSERVABLES;
        L.separate = flags & MFS_SEPARATE_OBSERVABLES;
        const uint32_t D = prog.num_detectors, O = prog.num_observables;
        L.ncols = D + O * ((L.prepend ? 1 : 0) + (L.append ? 1 : 0));
        L.det_row = L.packed ? (L.ncols + 7) / 8 : L.ncols;
        L.obs_row = L.packed ? (O + 7) / 8 : O;
        L.via_unpacked = L.packed && ((L.prepend && (O % 8)) || (L.append && (D % 8)));
        return L;
    }

    // Fills output descriptors for a batch written into device rows.
    void outputs(const Layout &L, uint8_t *det_dev, size_t det_stride, uint8_t *obs_dev, size_t obs_stride,
                 LaunchParams &P, char base_char = 0) {
        const uint32_t D = prog.num_detectors, O = prog.num_observables;
        const bool packed = L.packed && !L.via_unpacked;
        P.dets = DevOut{det_dev, det_stride, 0, packed ? 1u : 0u, (uint32_t)(uint8_t)base_char, 1};
        if (L.prepend) P.dets.col0 = O;
        if (L.prepend || L.append) {
            // Observables share the dets row. A second combined row copy goes to obs when separate.
            P.obs = DevOut{det_dev, det_stride, L.append ? D : 0u, packed ? 1u : 0u, (uint32_t)(uint8_t)base_char, 1};
        } else if (L.separate && obs_dev) {
            P.obs = DevOut{obs_dev, obs_stride, 0, L.packed ? 1u : 0u, (uint32_t)(uint8_t)base_char, 1};
        } else {
            P.obs = DevOut{nullptr, 0, 0, 0, 0, 0};
        }
    }

    template <class V>
    static void upload(DevBuf &b, const std::vector<V> &v, const char *what) {
        b.ensure(std::max<size_t>(v.size(), 1) * sizeof(V));
        if (!v.empty()) check(hipMemcpy(b.p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice), what);
    }
    void upload_program() {
        upload(d_blob, prog.blob, "upload program");
        upload(d_ntgt, prog.ntgt, "upload noise targets");
        upload(d_otab, prog.otab, "upload outcome table");
    }

    // A noiseless run with frame randomization: if every detector and observable stays zero over
    // many random frames, randomization cannot change any output (the outputs are linear in the
    // random bits), so sampling can skip it. Misses a nondeterministic output with probability 2^-512.
    void probe_randomization() {
        const uint64_t tiles = std::max<uint64_t>(1, 512 / W());
        const uint64_t shots = tiles * W();
        Layout L = layout(MFS_BIT_PACKED | MFS_SEPARATE_OBSERVABLES);
        d_out_det.ensure(shots * std::max<size_t>(L.det_row, 1));
        d_out_obs.ensure(shots * std::max<size_t>(L.obs_row, 1));
        check(hipMemsetAsync(d_out_det.p, 0, d_out_det.cap, stream), "memset");
        check(hipMemsetAsync(d_out_obs.p, 0, d_out_obs.cap, stream), "memset");
        LaunchParams P = base_params();
        outputs(L, d_out_det.as<uint8_t>(), L.det_row, d_out_obs.as<uint8_t>(), L.obs_row, P);
        P.seed = splitmix64(key ^ 0x5EEDull);
        P.shots = shots;
        P.tile0 = 0;
        P.mode = MODE_NOISELESS;
        P.randomize = 1;
        P.det_invert = nullptr;   // the probe looks at flips only
        launch(P, tiles);
        std::vector<uint8_t> a(shots * L.det_row), b(shots * L.obs_row);
        if (!a.empty()) check(hipMemcpyAsync(a.data(), d_out_det.p, a.size(), hipMemcpyDeviceToHost, stream), "copy");
        if (!b.empty()) check(hipMemcpyAsync(b.data(), d_out_obs.p, b.size(), hipMemcpyDeviceToHost, stream), "copy");
        check(hipStreamSynchronize(stream), "probe");
        randomize = false;
        for (uint8_t v : a) randomize |= v != 0;
        for (uint8_t v : b) randomize |= v != 0;
    }

    // Runs one sample call into host or device rows.
    void sample(uint64_t shots, uint32_t flags, uint8_t *dets, size_t dets_stride, uint8_t *obs, size_t obs_stride,
                const uint64_t *fault_offsets = nullptr, const uint32_t *faults = nullptr, char base_char = 0,
                bool newline = false, uint8_t *errs = nullptr, size_t errs_stride = 0) {
        Layout L = layout(flags);
        if (L.separate && (L.prepend || L.append)) {
            throw std::invalid_argument(
                "Can't specify separate_observables=True with append_observables=True or prepend_observables=True");
        }
        const bool device_out = flags & MFS_DEVICE_OUTPUT;
        if (shots == 0) return;
        if (use_expansion(faults != nullptr, shots)) {
            sample_expand(L, shots, device_out, dets, dets_stride, base_char, newline);
            return;
        }
        if (use_symptoms(faults != nullptr, shots)) {
            sample_symptom(L, shots, device_out, dets, dets_stride, obs, obs_stride, fault_offsets, faults, base_char,
                           newline, errs, errs_stride);
            return;
        }
        const uint64_t W64 = W();
        const size_t row_bytes = L.det_row + (L.separate ? L.obs_row : 0) + (L.via_unpacked ? L.ncols : 0);
        uint64_t batch = device_out && !L.via_unpacked ? shots
                                                        : std::max<uint64_t>(W64, (HOST_BATCH_BYTES / std::max<size_t>(row_bytes, 1)) / W64 * W64);
        if (device_out && L.via_unpacked) batch = arena_batch(batch, L.ncols, W64);
        const bool piped = !device_out;
        if (piped) {
            // Host rows travel through the pipeline: batches of about PIPE_BATCH_BYTES of host rows.
            const size_t host_row = std::max<size_t>(L.det_row + (L.separate ? L.obs_row : 0), 1);
            batch = std::max<uint64_t>(W64, PIPE_BATCH_BYTES / host_row / W64 * W64);
        }
        batch = std::min(batch, shots);
        PipeCall pipe_call{this, piped};
        if (piped) {
            const size_t dev_row = (L.via_unpacked ? L.ncols : 0) + std::max<size_t>(L.det_row, 1) +
                                   (L.separate ? std::max<size_t>(L.obs_row, 1) : 0);
            const size_t host_row = L.det_row + (L.separate ? L.obs_row : 0);
            pipe.begin(stream, batch * dev_row + 4 * 256, std::max<size_t>(batch * host_row, 256));
        }
        const bool fixed = faults != nullptr;
        const uint64_t call_tile0 = tiles_used;
        for (uint64_t s0 = 0; s0 < shots; s0 += batch) {
            const uint64_t n = std::min(batch, shots - s0);
            const uint64_t tiles = (n + W64 - 1) / W64;
            LaunchParams P = base_params();
            P.shots = n;
            P.tile0 = call_tile0 + s0 / W64;
            P.mode = fixed ? MODE_FIXED : MODE_SAMPLED;
            P.randomize = (!fixed && randomize) ? 1 : 0;
            if (fixed) stage_faults(s0, n, tiles, fault_offsets, faults, P);
            std::unique_ptr<SlotScope> slot;
            if (piped) slot.reset(new SlotScope(this));
            staging_reset();
            uint8_t *det_dst, *obs_dst;
            size_t det_stride_dev, obs_stride_dev;
            if (L.via_unpacked) {
                uint8_t *tmp = staging(d_tmp, n * L.ncols);
                Layout U = L;
                U.packed = false;
                U.via_unpacked = false;
                outputs(U, tmp, L.ncols, nullptr, 0, P);
                if (device_out) {
                    det_dst = dets + s0 * dets_stride;
                    det_stride_dev = dets_stride;
                } else {
                    det_dst = staging(d_out_det, n * L.det_row);
                    det_stride_dev = L.det_row;
                }
                launch(P, tiles);
                check(launch_pack_rows(tmp, L.ncols, det_dst, det_stride_dev, n, L.ncols, stream), "pack rows");
                obs_dst = nullptr;
                obs_stride_dev = 0;
            } else {
                if (device_out) {
                    det_dst = dets + s0 * dets_stride;
                    det_stride_dev = dets_stride;
                    obs_dst = (L.separate && obs) ? obs + s0 * obs_stride : nullptr;
                    obs_stride_dev = obs_stride;
                } else {
                    det_dst = staging(d_out_det, n * std::max<size_t>(L.det_row, 1));
                    det_stride_dev = L.det_row;
                    obs_dst = L.separate ? staging(d_out_obs, n * std::max<size_t>(L.obs_row, 1)) : nullptr;
                    obs_stride_dev = L.obs_row;
                }
                if (L.packed && (L.det_row * 8 != L.ncols)) {
                    // Padding bits of the last byte are written as zero by the flush of the final partial
                    // block; nothing to do. Rows with zero columns are skipped below.
                }
                outputs(L, det_dst, det_stride_dev, obs_dst, obs_stride_dev, P, base_char);
                if (L.ncols == 0 && !(L.separate && obs_dst)) {
                    // Nothing to write; still advance the RNG position.
                } else {
                    launch(P, tiles);
                }
            }
            if (piped) {
                std::vector<HostPipe::Part> parts;
                if (dets && (L.det_row || newline)) {
                    parts.push_back({det_dst, det_stride_dev, L.det_row, dets + s0 * dets_stride, dets_stride,
                                     newline ? (long)L.ncols : -1L});
                }
                if (L.separate && obs && L.obs_row && obs_dst) {
                    parts.push_back({obs_dst, obs_stride_dev, L.obs_row, obs + s0 * obs_stride, obs_stride, -1L});
                }
                pipe.submit(parts, n);
            }
        }
        if (piped) pipe.finish();
        check(hipStreamSynchronize(stream), "sample");
        tiles_used += (shots + W64 - 1) / W64;
    }

    // ------------------------------------------------------------------ measurement expansion
    // Plans the measurement sampler's expansion path (semantics in expand.h) and builds its detector
    // sampler and device tables; false (with expand_why) when it does not apply.
    bool ensure_expansion() {
        if (expand_state) return expand_state > 0;
        expand_state = -1;
        const uint32_t M = (uint32_t)prog.num_measurements;
        if (!meas_mode || M == 0) {
            expand_why = "no measurements to expand";
            return false;
        }
        if (m2d_lds_bytes((M + 31) / 32, 0) > MAX_LDS) {
            expand_why = "the expansion keeps a tile's measurement words in LDS: too many measurements";
            return false;
        }
        RefStats rs;
        try {
            (void)reference_sample(flat, &rs);
        } catch (const std::exception &e) {
            // A circuit too large for the host tableau keeps the frame engine.
            expand_why = e.what();
            return false;
        }
        auto d = std::make_unique<mfs_sampler>();
        d->flat = flat;
        d->prog = compile_program(flat);
        if (d->prog.num_measurements != M) {
            expand_why = "record bookkeeping mismatch";
            return false;
        }
        d->key = splitmix64(key ^ 0xD7EC7E5ull);
        d->stream = stream;
        d->borrowed_stream = true;
        d->pick_tile_words(0);
        d->upload_program();
        d->probe_randomization();
        if (d->randomize) {
            expand_why = "the circuit has random detectors or observables";
            return false;
        }
        if (!d->ensure_symptoms()) {
            expand_why = "symptom engine unavailable: " + d->sym_why;
            return false;
        }
        const uint32_t D = d->prog.num_detectors, O = d->prog.num_observables;
        const ExpandPlan E = plan_expansion(flat, M, D, O, rs.random_recorded, rs.random_total, EXPAND_THREADS);
        if (!E.complete) {
            expand_why = E.why;
            return false;
        }
        // The y rows are the detector sampler's packed rows (observables appended), padded to 16 bytes.
        d->row_pitch_override = std::max<size_t>(((D + O + 7) / 8 + 15) / 16 * 16, 16);
        auto nonempty = [](const std::vector<uint32_t> &v) { return v.empty() ? std::vector<uint32_t>{0, 0} : v; };
        upload(d_x_prim, nonempty(E.prim), "expansion y slots");
        upload(d_x_extra_off, E.extra_off, "expansion y slots");
        upload(d_x_extra, nonempty(E.extra), "expansion y slots");
        upload(d_x_free, nonempty(E.free_slots), "expansion free results");
        upload(d_x_sched, E.sched_off, "expansion schedule");
        upload(d_x_items, E.items.empty() ? std::vector<uint32_t>(EXPAND_RECORD_WORDS, 0) : E.items, "expansion items");
        upload(d_x_terms, nonempty(E.terms), "expansion terms");
        // Output inversion: bit m of the words is the reference bit of measurement m (zero past M).
        std::vector<uint32_t> inv((M + 31) / 32 + 1, 0);
        for (uint32_t m = 0; m < M; m++) {
            if (m / 32 < invert_words.size() && ((invert_words[m / 32] >> (m % 32)) & 1u)) inv[m / 32] |= 1u << (m % 32);
        }
        upload(d_x_inv, inv, "expansion output inversion");
        expand_free = (uint32_t)E.free_slots.size();
        expand_phases = E.phases;
        expand_items = (uint32_t)(E.items.size() / EXPAND_RECORD_WORDS);
        expand_key = splitmix64(key ^ 0xE8A7D5ull);
        det_src = std::move(d);
        expand_state = 1;
        return true;
    }

    // Whether a measurement-sampler call expands detection events: never with fixed faults or on the
    // frame engine setting; automatically once the detector sampler's table would cost at most two
    // columns per shot of the call (as the symptom engine decides), or when already built.
    bool use_expansion(bool fixed, uint64_t shots) {
        if (!meas_mode || fixed || engine_mode == 1) return false;
        if (engine_mode == 2) {
            if (arena && expand_state == 0)
                throw std::invalid_argument("the sampler's expansion is not built yet (building it allocates device "
                                            "memory); prepare the sampler first with mfs_engine_info or one allocating call");
            if (!ensure_expansion()) throw std::invalid_argument("expansion unavailable: " + expand_why);
            return true;
        }
        if (expand_state < 0) return false;
        if (expand_state == 0) {
            if (shots != ANY_SHOTS) {
                if (!plan_counted) {
                    column_count = plan_symptoms(compile_program(flat)).columns.size();
                    plan_counted = true;
                }
                if (column_count > 2 * shots) return false;
            }
            if (arena)
                throw std::invalid_argument("the sampler's expansion is not built yet (building it allocates device "
                                            "memory); prepare the sampler first with mfs_engine_info or one allocating call");
        }
        return ensure_expansion();
    }

    ExpandDev expand_params(uint64_t n, const uint8_t *y, uint8_t *out, size_t out_stride, uint64_t tile0) {
        const mfs_sampler &d = *det_src;
        const uint32_t M = (uint32_t)prog.num_measurements;
        ExpandDev E{};
        E.in = y;
        E.in_stride = d.row_pitch_override;
        E.in_groups = (uint32_t)(d.row_pitch_override / 16);
        E.Y = d.prog.num_detectors + d.prog.num_observables;
        E.prim = d_x_prim.as<uint32_t>();
        E.extra_off = d_x_extra_off.as<uint32_t>();
        E.extra = d_x_extra.as<uint32_t>();
        E.free_slots = d_x_free.as<uint32_t>();
        E.nfree = expand_free;
        E.nphases = expand_phases;
        E.sched_off = d_x_sched.as<uint32_t>();
        E.items = d_x_items.as<uint4>();
        E.terms = d_x_terms.as<uint32_t>();
        E.seed = expand_key;
        E.tile0 = tile0;
        M2dParams &P = E.out;
        P.blocks = (M + 31) / 32;
        P.num_det = M;
        P.out = out;
        P.out_stride = out_stride;
        P.n_out = M;
        P.n_obs = 0;
        P.shots = n;
        E.inv = d_x_inv.as<uint32_t>();
        return E;
    }

    // Measurement records of `shots` shots: detection events and observable flips on the detector
    // sampler (padded y rows in device staging), then the expansion kernel into packed rows; bool or
    // text outputs are expanded from those. Batches are whole tiles of both streams (the detector
    // sampler's symptom tiles and the expansion's 32-shot tiles).
    void sample_expand(const Layout &L, uint64_t shots, bool device_out, uint8_t *out, size_t out_stride,
                       char base_char, bool newline) {
        mfs_sampler &d = *det_src;
        const uint32_t M = (uint32_t)prog.num_measurements;
        const size_t ypitch = d.row_pitch_override, prow = (M + 7) / 8;
        const Layout DL = d.layout(MFS_BIT_PACKED | MFS_APPEND_OBSERVABLES);
        static uint8_t dummy;
        const uint64_t S = d.sym_prepare(DL, 1, &dummy, ypitch, nullptr, 0, 0).P.shots_per_tile;
        const uint64_t unit = std::max<uint64_t>(S, 32);
        const bool piped = !device_out;
        const size_t host_row = L.packed ? prow : M;
        uint64_t batch = std::max<uint64_t>(unit, HOST_BATCH_BYTES / (ypitch + prow) / unit * unit);
        if (piped) batch = std::max<uint64_t>(unit, PIPE_BATCH_BYTES / std::max<size_t>(host_row, 1) / unit * unit);
        if (device_out) batch = arena_batch(batch, ypitch + (L.packed ? 0 : prow), unit);
        batch = std::min(batch, shots);
        PipeCall pipe_call{this, piped};
        if (piped) {
            const size_t dev_row = ypitch + prow + (L.packed ? 0 : M);
            pipe.begin(stream, batch * dev_row + 6 * 256, std::max<size_t>(batch * host_row, 256));
        }
        const uint64_t d_tile0 = d.tiles_used, x_tile0 = expand_tiles;
        uint64_t d_tiles = 0;
        for (uint64_t s0 = 0; s0 < shots; s0 += batch) {
            const uint64_t n = std::min(batch, shots - s0);
            std::unique_ptr<SlotScope> slot;
            if (piped) slot.reset(new SlotScope(this));
            staging_reset();
            uint8_t *y = staging(d_x_y, n * ypitch);
            d_tiles += d.sym_launch(DL, n, y, ypitch, nullptr, 0, d_tile0 + d_tiles, nullptr, nullptr);
            uint8_t *pd;
            size_t pd_stride;
            if (L.packed && device_out) {
                pd = out + s0 * out_stride;
                pd_stride = out_stride;
            } else {
                pd = staging(d_out_det, n * prow);
                pd_stride = prow;
            }
            check(launch_expand(expand_params(n, y, pd, pd_stride, x_tile0 + s0 / 32), stream), "expansion launch");
            if (!L.packed) {
                uint8_t *ud = device_out ? out + s0 * out_stride : staging(d_tmp, n * std::max<size_t>(M, 1));
                const size_t ud_stride = device_out ? out_stride : M;
                check(launch_unpack_rows(pd, pd_stride, ud, ud_stride, n, M, (uint8_t)base_char, stream), "unpack");
                pd = ud;
                pd_stride = ud_stride;
            }
            if (piped) {
                std::vector<HostPipe::Part> parts;
                if (out) parts.push_back({pd, pd_stride, host_row, out + s0 * out_stride, out_stride,
                                          newline ? (long)M : -1L});
                pipe.submit(parts, n);
            }
        }
        if (piped) pipe.finish();
        check(hipStreamSynchronize(stream), "sample");
        d.tiles_used += d_tiles;
        expand_tiles += (shots + 31) / 32;
    }

    // ------------------------------------------------------------------ symptom engine
    // Builds the symptom table: one fixed-fault frame-kernel shot per basis fault, then the set
    // bits of each shot's detector/observable row become that column's sparse symptom.
    bool ensure_symptoms() {
        if (sym_state) return sym_state > 0;
        plan = plan_symptoms(prog);
        if (!plan.supported) {
            sym_state = -1;
            sym_why = plan.why_unsupported;
            return false;
        }
        const uint64_t F = plan.columns.size();
        Layout L = layout(MFS_BIT_PACKED | MFS_SEPARATE_OBSERVABLES);
        const size_t drow = std::max<size_t>(L.det_row, 1), orow = std::max<size_t>(L.obs_row, 1);
        const uint64_t W64 = W();
        uint64_t batch = std::max<uint64_t>(W64, ((size_t)2 << 30) / (drow + orow) / W64 * W64);
        std::vector<uint32_t> offs(F + 1, 0);
        std::vector<uint32_t> ents;
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "event");
        check(hipEventCreate(&e1), "event");
        check(hipEventRecord(e0, stream), "event");
        for (uint64_t c0 = 0; c0 < F; c0 += batch) {
            const uint64_t n = std::min(batch, F - c0);
            const uint64_t tiles = (n + W64 - 1) / W64;
            // Fixed faults: shot i of the batch carries column c0 + i's basis fault.
            std::vector<DevFault> fs(n);
            std::vector<uint32_t> fo(tiles + 1, 0);
            for (uint64_t t = 0; t < tiles; t++) {
                fo[t] = (uint32_t)(t * W64);
                const uint64_t e = std::min(n, (t + 1) * W64);
                for (uint64_t i = t * W64; i < e; i++) {
                    const SymColumn &cdef = plan.columns[c0 + i];
                    fs[i] = DevFault{cdef.op, cdef.site, (uint32_t)(i - t * W64), cdef.code};
                }
                std::stable_sort(fs.begin() + (long)(t * W64), fs.begin() + (long)e,
                                 [](const DevFault &a, const DevFault &b) { return a.op < b.op; });
            }
            fo[tiles] = (uint32_t)n;
            d_faults.ensure(fs.size() * sizeof(DevFault));
            d_fault_off.ensure(fo.size() * 4);
            check(hipMemcpyAsync(d_faults.p, fs.data(), fs.size() * sizeof(DevFault), hipMemcpyHostToDevice, stream), "faults");
            check(hipMemcpyAsync(d_fault_off.p, fo.data(), fo.size() * 4, hipMemcpyHostToDevice, stream), "fault offsets");
            d_out_det.ensure(n * drow);
            d_out_obs.ensure(n * orow);
            LaunchParams P = base_params();
            P.det_invert = nullptr;   // symptom columns are flips only
            outputs(L, d_out_det.as<uint8_t>(), drow, d_out_obs.as<uint8_t>(), orow, P);
            P.shots = n;
            P.tile0 = 0;
            P.mode = MODE_FIXED;
            P.randomize = 0;
            P.faults = d_faults.as<DevFault>();
            P.fault_off = d_fault_off.as<uint32_t>();
            if (L.det_row == 0) P.dets.enabled = 0;
            if (L.obs_row == 0) P.obs.enabled = 0;
            launch(P, tiles);
            d_sym_tmp.ensure((n + 1) * 4);
            check(launch_sym_count(d_out_det.as<uint8_t>(), (uint32_t)L.det_row, d_out_obs.as<uint8_t>(),
                                   (uint32_t)L.obs_row, n, d_sym_tmp.as<uint32_t>(), stream), "symptom count");
            std::vector<uint32_t> counts(n);
            check(hipMemcpyAsync(counts.data(), d_sym_tmp.p, n * 4, hipMemcpyDeviceToHost, stream), "counts");
            check(hipStreamSynchronize(stream), "symptom count");
            std::vector<uint32_t> local(n + 1, 0);
            for (uint64_t i = 0; i < n; i++) local[i + 1] = local[i] + counts[i];
            for (uint64_t i = 0; i < n; i++) offs[c0 + i + 1] = offs[c0 + i] + counts[i];
            d_sym_tmp.ensure((n + 1) * 4);
            check(hipMemcpyAsync(d_sym_tmp.p, local.data(), (n + 1) * 4, hipMemcpyHostToDevice, stream), "offsets");
            DevBuf bent;
            bent.ensure(std::max<size_t>(local[n], 1) * 4);
            check(launch_sym_fill(d_out_det.as<uint8_t>(), (uint32_t)L.det_row, d_out_obs.as<uint8_t>(),
                                  (uint32_t)L.obs_row, n, prog.num_detectors, d_sym_tmp.as<uint32_t>(),
                                  bent.as<uint32_t>(), stream), "symptom fill");
            const size_t base = ents.size();
            ents.resize(base + local[n]);
            if (local[n]) check(hipMemcpyAsync(ents.data() + base, bent.p, local[n] * 4, hipMemcpyDeviceToHost, stream), "ents");
            check(hipStreamSynchronize(stream), "symptom fill");
        }
        check(hipEventRecord(e1, stream), "event");
        check(hipEventSynchronize(e1), "event");
        float ms = 0;
        check(hipEventElapsedTime(&ms, e0, e1), "event");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        sym_build_ms = ms;
        sym_entries = ents.size();
        upload(d_classes, plan.classes, "classes");
        upload(d_sotab, plan.otab, "symptom outcome table");
        // Column records: up to two entries inline, longer symptoms in the overflow list.
        std::vector<uint2> recs(F);
        std::vector<uint32_t> ovf;
        for (uint64_t c = 0; c < F; c++) {
            const uint32_t n = offs[c + 1] - offs[c];
            const uint32_t *e = ents.data() + offs[c];
            if (n <= SYM_INLINE_MAX) {
                recs[c] = make_uint2((n << 28) | (n >= 1 ? e[0] : 0u), n == 2 ? e[1] : 0u);
            } else {
                if (ovf.size() >= (1u << 28)) throw std::invalid_argument("symptom table too large");
                recs[c] = make_uint2((3u << 28) | (uint32_t)ovf.size(), n);
                ovf.insert(ovf.end(), e, e + n);
            }
        }
        if (prog.num_detectors + prog.num_observables >= (1u << 28)) {
            sym_state = -1;
            sym_why = "too many detectors for symptom records";
            return false;
        }
        upload(d_recs, recs, "symptom records");
        upload(d_ovf, ovf, "symptom overflow entries");
        // Compact records by slot, when every entry index fits 15 bits.
        sym_compact = prog.num_detectors <= SYM_COMPACT_DETECTORS && prog.num_observables <= SYM_COMPACT_OBSERVABLES;
        if (sym_compact) {
            std::vector<uint32_t> r32(std::max<uint64_t>(plan.total_slots, 1), 0);
            std::vector<uint32_t> o32;
            for (const SymClass &cl : plan.classes) {
                for (uint64_t st = 0; st < cl.nsites; st++) {
                    for (uint32_t j = 0; j < cl.basis; j++) {
                        const uint64_t c = cl.col0 + st * cl.basis + j;
                        const uint32_t n = offs[c + 1] - offs[c];
                        const uint32_t *e = ents.data() + offs[c];
                        uint32_t dets[2], nd = 0, obs_bits = 0;
                        bool spill = false;
                        for (uint32_t t = 0; t < n; t++) {
                            if (e[t] >= prog.num_detectors) {
                                obs_bits |= 1u << (30 + (e[t] - prog.num_detectors));
                            } else if (nd < 2) {
                                dets[nd++] = e[t];
                            } else {
                                spill = true;
                            }
                        }
                        uint32_t rec;
                        if (!spill) {
                            rec = obs_bits | (nd << 28) | (nd >= 1 ? dets[0] : 0u) | (nd == 2 ? dets[1] << 14 : 0u);
                        } else {
                            if (o32.size() >= (1u << 28)) throw std::invalid_argument("symptom table too large");
                            rec = (3u << 28) | (uint32_t)o32.size();
                            o32.push_back(n);
                            o32.insert(o32.end(), e, e + n);
                        }
                        r32[cl.rec0 + st * cl.rstride + j] = rec;
                    }
                }
            }
            sym_ovf32 = !o32.empty();
            upload(d_recs32, r32, "compact symptom records");
            upload(d_ovf32, o32, "compact symptom overflow entries");
            build_outcome_records(offs, ents);
        } else {
            // Without compact records a hit needs its class's column layout: one group per class.
            plan.groups.clear();
            for (uint32_t c = 0; c < plan.classes.size(); c++) {
                const SymClass &sc = plan.classes[c];
                SymGroup g{};
                g.p = sc.p;
                g.inv_log = sc.inv_log;
                g.site0 = sc.site0;
                g.nsites = sc.nsites;
                g.rec0 = sc.rec0;
                g.rstride = sc.rstride;
                g.nmem = 1;
                g.c0 = c;
                for (int k = 0; k < 4; k++) g.bound[k] = 0xFFFFFFFFu;
                g.bound[0] = 0;
                g.nout[0] = sc.nout;
                g.woff[0] = sc.woff;
                plan.groups.push_back(g);
            }
        }
        upload(d_groups, plan.groups, "symptom groups");
        sym_state = 1;
        return true;
    }

    // Outcome records for the specialised scan kernel: when every sampling group is a plain one (compact
    // records, four slots per site, no thinning) whose members draw outcomes uniformly over codes
    // 1..nout, each (site, code) gets the symptom of that outcome precomputed: the XOR of the basis
    // columns the code selects (bit j = column j, as the generic kernel applies them), with entries
    // that appear an even number of times removed. A hit then reads its site's entry and one record
    // instead of four basis records. Each site's block of nout records is written relative to the
    // site's lowest detector and stored once per distinct block, so that the table stays small enough
    // to live in L2 (a d=25 memory circuit's 980,000 outcomes reduce to the blocks of about three rounds).
    void build_outcome_records(const std::vector<uint32_t> &offs, const std::vector<uint32_t> &ents) {
        sym_fast = false;
        const uint32_t D = prog.num_detectors;
        for (const SymGroup &g : plan.groups) {
            if (g.flags != 0 || g.rstride != 4) return;
            for (uint32_t k = 0; k < g.nmem; k++)
                if (!(g.nout[k] & SYM_SEQ) || (g.nout[k] & 0xFFFFu) > 15) return;
        }
        std::vector<uint64_t> site0(plan.classes.size());
        uint64_t nsites = 0;
        for (size_t c = 0; c < plan.classes.size(); c++) {
            site0[c] = nsites;
            nsites += plan.classes[c].nsites;
        }
        if (nsites >= (1ull << 31)) return;
        std::vector<uint2> sites(std::max<uint64_t>(nsites, 1), make_uint2(0, 0));
        std::vector<uint2> orec;
        std::vector<uint32_t> oovf;
        std::vector<uint32_t> acc;
        // Distinct relative blocks: key = nout, then per code its relative entries and a terminator.
        std::unordered_map<std::string, uint32_t> blocks;
        std::vector<std::vector<uint32_t>> site_syms;
        std::string key;
        for (size_t c = 0; c < plan.classes.size(); c++) {
            const SymClass &cl = plan.classes[c];
            const uint32_t nout = cl.nout & 0xFFFFu;
            site_syms.resize(nout);
            for (uint64_t st = 0; st < cl.nsites; st++) {
                uint32_t lo = D;
                for (uint32_t code = 1; code <= nout; code++) {
                    acc.clear();
                    for (uint32_t j = 0; j < cl.basis; j++) {
                        if (!((code >> j) & 1)) continue;
                        const uint64_t col = cl.col0 + st * cl.basis + j;
                        acc.insert(acc.end(), ents.begin() + offs[col], ents.begin() + offs[col + 1]);
                    }
                    std::sort(acc.begin(), acc.end());
                    size_t w = 0;
                    for (size_t i = 0; i < acc.size();) {
                        if (i + 1 < acc.size() && acc[i] == acc[i + 1]) {
                            i += 2;
                        } else {
                            acc[w++] = acc[i++];
                        }
                    }
                    acc.resize(w);
                    for (uint32_t e : acc)
                        if (e < D) lo = std::min(lo, e);
                    site_syms[code - 1] = acc;
                }
                const uint32_t b = lo < D ? lo : 0;
                key.assign(reinterpret_cast<const char *>(&nout), 4);
                for (uint32_t code = 1; code <= nout; code++) {
                    for (uint32_t e : site_syms[code - 1]) {
                        const uint32_t v = e < D ? e - b : e;
                        key.append(reinterpret_cast<const char *>(&v), 4);
                    }
                    const uint32_t end = 0xFFFFFFFFu;
                    key.append(reinterpret_cast<const char *>(&end), 4);
                }
                auto found = blocks.find(key);
                if (found != blocks.end()) {
                    sites[site0[c] + st] = make_uint2(found->second | (nout << 28), b);
                    continue;
                }
                const uint64_t first = orec.size();
                if (first + nout >= (1ull << 28) || (first + nout) * 8 > ((uint64_t)1 << 30)) return;
                blocks.emplace(key, (uint32_t)first);
                sites[site0[c] + st] = make_uint2((uint32_t)first | (nout << 28), b);
                for (uint32_t code = 1; code <= nout; code++) {
                    acc = site_syms[code - 1];
                    for (uint32_t &e : acc)
                        if (e < D) e -= b;
                    uint32_t dets[4], nd = 0, obs = 0;
                    bool spill = false;
                    for (uint32_t e : acc) {
                        if (e >= D) obs |= 1u << (e - D);
                        else if (nd < 4) dets[nd++] = e;
                        else spill = true;
                    }
                    uint2 r;
                    if (!spill) {
                        for (uint32_t i = nd; i < 4; i++) dets[i] = 0;
                        r.x = dets[0] | (dets[1] << 14) | (((1u << nd) - 1u) << 28);
                        r.y = dets[2] | (dets[3] << 14) | (((obs >> 1) & 1u) << 28) | ((obs & 1u) << 29);
                    } else {
                        if (oovf.size() >= (1u << 30)) return;
                        r = make_uint2(0, SYM_OREC_OVERFLOW | (uint32_t)oovf.size());
                        oovf.push_back((uint32_t)acc.size());
                        oovf.insert(oovf.end(), acc.begin(), acc.end());
                    }
                    orec.push_back(r);
                }
            }
        }
        std::vector<uint32_t> base(4 * std::max<size_t>(plan.groups.size(), 1), 0);
        for (size_t gi = 0; gi < plan.groups.size(); gi++) {
            const SymGroup &g = plan.groups[gi];
            // Member k's sites follow member k - 1's in the site table, as in the group's site order.
            for (uint32_t k = 0; k < g.nmem; k++) {
                base[4 * gi + k] = (uint32_t)(site0[g.c0 + k] - g.bound[k]);
                if (base[4 * gi + k] != base[4 * gi]) return;
            }
        }
        if (orec.empty()) orec.push_back(make_uint2(0, 0));
        upload(d_orecs, orec, "outcome records");
        upload(d_orec_base, base, "outcome record site bases");
        upload(d_orec_sites, sites, "outcome record sites");
        upload(d_orec_ovf, oovf, "outcome record overflow lists");
        orec_blocks = blocks.size();
        orec_records = orec.size();
        sym_fast = true;
    }

    // Builds the symptom tables of a detector error model (reference DemSampler,
    // src/stim/simulators/dem_sampler.inl:50-77: each mechanism fires independently with its
    // probability and XORs its targets into the outputs). Mechanisms are ordered by probability and
    // cut into groups whose probabilities lie within a factor DEM_BUCKET_RATIO of the group's largest;
    // each group is drawn at that largest probability and thinned per mechanism (SYM_THIN).
    static constexpr double DEM_BUCKET_RATIO = 0.5;
    void init_dem(const DemModel &m) {
        dem_mode = true;
        engine_mode = 2;
        randomize = false;
        if (m.num_detectors >= (1ull << 31) || m.num_observables >= (1ull << 31))
            throw std::invalid_argument("detector error model too large");
        prog.num_detectors = (uint32_t)m.num_detectors;
        prog.num_observables = (uint32_t)m.num_observables;
        const uint64_t E = m.errors.size();
        if (E >= (1ull << 31)) throw std::invalid_argument("too many error mechanisms");
        std::vector<uint32_t> order(E);
        for (uint32_t e = 0; e < E; e++) order[e] = e;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return m.errors[a].p > m.errors[b].p; });
        plan = SymPlan{};
        dem_col.assign(E, 0);
        sym_compact = prog.num_detectors <= SYM_COMPACT_DETECTORS && prog.num_observables <= SYM_COMPACT_OBSERVABLES;
        const uint32_t D = prog.num_detectors;
        std::vector<uint2> recs(E);
        std::vector<uint32_t> ovf, r32, o32, thr(std::max<uint64_t>(E, 1), 0xFFFFFFFFu);
        // Outcome table shared by every class: one outcome, code 1.
        plan.otab = {0xFFFFFFFFu, 1u};
        uint64_t slot = 0;
        for (uint64_t i = 0; i < E;) {
            const double pmax = m.errors[order[i]].p;
            uint64_t e = i;
            auto joins = [&](double p) { return p == pmax || (pmax > 0 && p > pmax * DEM_BUCKET_RATIO); };
            while (e < E && joins(m.errors[order[e]].p)) e++;
            SymClass sc{};
            sc.p = (float)pmax;
            sc.inv_log = (pmax > 0 && pmax < 1) ? (float)(1.0 / std::log2(1.0 - pmax)) : 0.0f;
            sc.nsites = (uint32_t)(e - i);
            sc.basis = 1;
            sc.col0 = (uint32_t)i;
            sc.woff = 0;
            sc.nout = 1u | SYM_UNIFORM | SYM_SEQ;
            sc.site0 = (uint32_t)i;
            sc.rec0 = (uint32_t)slot;
            sc.rstride = 4;
            SymGroup g{};
            g.p = sc.p;
            g.inv_log = sc.inv_log;
            g.site0 = sc.site0;
            g.nsites = sc.nsites;
            g.rec0 = sc.rec0;
            g.rstride = 4;
            g.nmem = 1;
            g.c0 = (uint32_t)plan.classes.size();
            for (int k = 0; k < 4; k++) g.bound[k] = 0xFFFFFFFFu;
            g.bound[0] = 0;
            g.nout[0] = sc.nout;
            g.woff[0] = 0;
            if (sym_compact) g.flags = SYM_DEM_SLOTS;
            for (uint64_t j = i; j < e; j++) {
                const DemError &err = m.errors[order[j]];
                dem_col[order[j]] = (uint32_t)j;
                plan.columns.push_back({order[j], 0, 1});
                if (pmax > 0) {
                    const double t = std::floor(err.p / pmax * 4294967296.0 + 0.5) - 1.0;
                    thr[j] = t >= 4294967295.0 ? 0xFFFFFFFFu : t < 0 ? 0u : (uint32_t)t;
                    if (thr[j] != 0xFFFFFFFFu) g.flags |= SYM_THIN;
                }
            }
            plan.classes.push_back(sc);
            plan.groups.push_back(g);
            slot += sc.nsites * 4ull;
            i = e;
        }
        plan.total_sites = E;
        plan.total_slots = slot;
        // Records: 8-byte per column (entries: detectors, then observables as D + k), compact by slot.
        if (sym_compact) r32.assign(std::max<uint64_t>(slot, 1), 0);
        uint32_t c = 0;   // class of mechanism j (classes cover increasing, contiguous site ranges)
        for (uint64_t j = 0; j < E; j++) {
            const DemError &err = m.errors[plan.columns[j].op];
            std::vector<uint32_t> ent(err.dets.begin(), err.dets.end());
            for (uint32_t k : err.obs) ent.push_back(D + k);
            const uint32_t n = (uint32_t)ent.size();
            if (n <= SYM_INLINE_MAX) {
                recs[j] = make_uint2((n << 28) | (n >= 1 ? ent[0] : 0u), n == 2 ? ent[1] : 0u);
            } else {
                if (ovf.size() >= (1u << 28)) throw std::invalid_argument("detector error model too large");
                recs[j] = make_uint2((3u << 28) | (uint32_t)ovf.size(), n);
                ovf.insert(ovf.end(), ent.begin(), ent.end());
            }
            if (sym_compact) {
                // Up to six detectors inline, two per record in slots 0, 2 and 3 (SYM_DEM_SLOTS).
                const size_t nd = err.dets.size();
                auto pair = [&](size_t a) -> uint32_t {
                    const size_t c = nd > a ? std::min<size_t>(2, nd - a) : 0;
                    return ((uint32_t)c << 28) | (c >= 1 ? err.dets[a] : 0u) | (c == 2 ? err.dets[a + 1] << 14 : 0u);
                };
                uint32_t rec, rec2 = 0, rec3 = 0;
                if (nd <= 6) {
                    rec = pair(0);
                    rec2 = pair(2);
                    rec3 = pair(4);
                    for (uint32_t k : err.obs) rec |= 1u << (30 + k);
                } else {
                    rec = (3u << 28) | (uint32_t)o32.size();
                    o32.push_back(n);
                    o32.insert(o32.end(), ent.begin(), ent.end());
                }
                // Mechanism j is site j: its class's records start at rec0 with 4 slots per site.
                while (c + 1 < plan.classes.size() && plan.classes[c + 1].site0 <= j) c++;
                const uint64_t at = plan.classes[c].rec0 + (j - plan.classes[c].site0) * 4ull;
                r32[at] = rec;
                r32[at + 1] = thr[j];
                r32[at + 2] = rec2;
                r32[at + 3] = rec3;
            }
        }
        sym_entries = 0;
        for (const DemError &err : m.errors) sym_entries += err.dets.size() + err.obs.size();
        upload(d_classes, plan.classes, "classes");
        upload(d_sotab, plan.otab, "symptom outcome table");
        upload(d_recs, recs, "symptom records");
        upload(d_ovf, ovf, "symptom overflow entries");
        upload(d_groups, plan.groups, "symptom groups");
        upload(d_thr, thr, "thinning thresholds");
        {
            std::vector<uint32_t> site_mech(std::max<uint64_t>(E, 1), 0);
            for (uint64_t j = 0; j < E; j++) site_mech[j] = plan.columns[j].op;
            upload(d_site_mech, site_mech, "mechanism of each site");
        }
        if (sym_compact) {
            sym_ovf32 = !o32.empty();
            upload(d_recs32, r32, "compact symptom records");
            upload(d_ovf32, o32, "compact symptom overflow entries");
        }
        sym_state = 1;
    }

    // Whether a call of `shots` shots runs on the symptom engine. In automatic mode the symptom table
    // is built only when it is small next to the work it saves: its build simulates one shot per basis
    // fault, so a call of fewer than half as many shots as there are columns samples on the frame
    // engine instead (until a larger call builds the table). ANY_SHOTS skips that test.
    static constexpr uint64_t ANY_SHOTS = ~0ull;
    bool use_symptoms(bool fixed, uint64_t shots = ANY_SHOTS) {
        if (engine_mode == 1 || meas_mode) return false;
        if ((prog.num_detectors + 2ull * prog.num_observables) / 8 > 64 * 1024) {
            if (engine_mode == 2) throw std::invalid_argument("rows too wide for the symptom engine's LDS image");
            return false;
        }
        if (fixed && engine_mode != 2) return false;
        if (randomize) {
            if (engine_mode == 2) throw std::invalid_argument(
                "the symptom engine needs deterministic detectors and observables; this circuit has random ones");
            return false;
        }
        if (engine_mode == 0 && sym_state == 0 && shots != ANY_SHOTS) {
            if (!plan_counted) {
                column_count = plan_symptoms(prog).columns.size();
                plan_counted = true;
            }
            if (column_count > 2 * shots) return false;
        }
        if (arena && sym_state == 0)
            throw std::invalid_argument(
                "the sampler's symptom table is not built yet (building it allocates device memory); prepare the "
                "sampler first with mfs_engine_info or one allocating call");
        const bool ok = ensure_symptoms();
        if (!ok && engine_mode == 2) throw std::invalid_argument("symptom engine unavailable: " + sym_why);
        return ok;
    }

    // Shots per symptom tile: the largest power of two (up to max_shots) whose output rows fit about
    // kb KB of LDS.
    uint32_t sym_tile_shots(size_t row_bytes, int kb, uint32_t max_shots) const {
        uint32_t S = 1;
        while (S < max_shots && (size_t)(2 * S) * row_bytes <= (size_t)kb * 1024) S *= 2;
        return S;
    }

    SymDevParams sym_params(uint32_t S) {
        SymDevParams P{};
        P.classes = d_classes.as<SymClass>();
        P.groups = d_groups.as<SymGroup>();
        P.ngroups = (uint32_t)plan.groups.size();
        P.otab = d_sotab.as<uint32_t>();
        P.recs = d_recs.as<uint2>();
        P.ovf = d_ovf.as<uint32_t>();
        P.recs32 = sym_compact ? d_recs32.as<uint32_t>() : nullptr;
        P.ovf32 = sym_compact ? d_ovf32.as<uint32_t>() : nullptr;
        P.ovf32_any = sym_ovf32 ? 1u : 0u;
        P.thr = d_thr.as<uint32_t>();
        P.nclasses = (uint32_t)plan.classes.size();
        P.otab_words = (uint32_t)plan.otab.size();
        // DEM samplers have many probability buckets per tile: each wave draws its share of them.
        P.wave_split = dem_mode ? 1u : 0u;
        P.D = prog.num_detectors;
        P.O = prog.num_observables;
        P.shots_per_tile = S;
        uint32_t lg = 0;
        while ((1u << lg) < S) lg++;
        P.tile_log2 = lg;
        P.total_sites = plan.total_sites;
        P.seed = splitmix64(key ^ 0x5A5A5A5Aull);
        P.orecs = nullptr;
        P.orec_base = nullptr;
        P.orec_sites = nullptr;
        P.orec_ovf = nullptr;
        P.err_words = err_words_cur;
        P.err_stride_words = err_stride_cur;
        P.site_mech = dem_mode ? d_site_mech.as<uint32_t>() : nullptr;
        return P;
    }

    // Sampled (or fixed-fault) shots through the symptom engine into packed device rows.
    // Kernel parameters of one symptom-engine launch over n shots starting at RNG tile tile0, with
    // tiles of about tile_kb KB of output rows (0: this sampler's setting).
    struct SymLaunch {
        SymDevParams P;
        uint64_t tiles;
        size_t lds_scan;     // dynamic LDS of the scan kernel
        size_t lds_queue;    // dynamic LDS of the queue and fixed-fault kernels
        bool scan_ok;        // tile positions fit the scan kernel's 31-bit range
        bool fast;           // the specialised scan kernel applies (outcome records built)
    };

    // The scan kernel takes tiles of up to 8192 shots (rows of a few bytes); the queue and
    // fixed-fault kernels take up to 1024.
    SymLaunch sym_prepare(const Layout &L, uint64_t n, uint8_t *det_dst, size_t det_stride, uint8_t *obs_dst,
                          size_t obs_stride, uint64_t tile0, int tile_kb = 0, bool scan = true) {
        const uint32_t D = prog.num_detectors, O = prog.num_observables;
        const size_t obs_bytes = (L.separate && obs_dst) ? (O + 7) / 8 : 0;
        const size_t det_bytes = row_pitch_override ? row_pitch_override : (L.ncols + 7) / 8;
        const size_t row = std::max<size_t>(det_bytes + obs_bytes, 1);
        const int kb = tile_kb ? tile_kb : sym_tile_kb;
        uint32_t S = sym_tile_shots(row, kb, scan ? 8192u : 1024u);
        auto log2_of = [](uint32_t v) {
            uint32_t lg = 0;
            while ((1u << lg) < v) lg++;
            return lg;
        };
        // The scan kernel addresses (site, shot) positions of one group in 32 bits.
        uint64_t group_sites = 0;
        for (const SymGroup &g : plan.groups) group_sites = std::max<uint64_t>(group_sites, g.nsites);
        if (dem_mode && scan) {
            // Thinned buckets are sampled only by the scan kernel: narrow the tile until it fits.
            while (S > 1 && (group_sites << log2_of(S)) >= (1ull << 31)) S /= 2;
        }
        bool scan_ok = scan && (group_sites << log2_of(S)) < (1ull << 31);
        if (dem_mode && scan && !scan_ok) throw std::invalid_argument("detector error model too large to sample");
        if (!scan_ok) S = std::min<uint32_t>(S, 1024);
        SymLaunch r{};
        SymDevParams &P = r.P;
        P = sym_params(S);
        P.shots = n;
        P.tile0 = tile0;
        P.dets = DevOut{det_dst, det_stride, L.prepend ? O : 0u, 1, 0, det_bytes ? 1u : 0u};
        P.det_row_bytes = (uint32_t)det_bytes;
        P.obs_mode = (L.prepend || L.append) ? 2u : obs_bytes ? 1u : 0u;
        P.obs_col0 = L.append ? D : 0;
        P.obs = DevOut{obs_dst, obs_stride, 0, 1, 0, obs_bytes ? 1u : 0u};
        P.obs_row_bytes = (uint32_t)obs_bytes;
        const size_t image = ((size_t)S * det_bytes + 15) / 16 * 16 + ((size_t)S * obs_bytes + 15) / 16 * 16;
        const size_t otab_bytes = P.otab_words <= SYM_OTAB_LDS_WORDS ? (size_t)P.otab_words * 4 : 0;
        r.lds_queue = image + SYM_QUEUE_BYTES + otab_bytes;
        r.lds_scan = image + 32 + otab_bytes;
        r.tiles = (n + S - 1) / S;
        if (r.tiles >= (1ull << 31)) throw std::invalid_argument("too many shots in one launch");
        r.scan_ok = scan_ok;
        r.fast = scan_ok && sym_fast && sym_variant == 0 && !P.err_words;
        if (r.fast) {
            P.orecs = d_orecs.as<uint2>();
            P.orec_base = d_orec_base.as<uint32_t>();
            P.orec_sites = d_orec_sites.as<uint2>();
            P.orec_ovf = d_orec_ovf.as<uint32_t>();
        }
        return r;
    }

    // Launches the symptom engine over n shots starting at RNG tile tile0; returns the tiles used.
    uint64_t sym_launch(const Layout &L, uint64_t n, uint8_t *det_dst, size_t det_stride, uint8_t *obs_dst,
                        size_t obs_stride, uint64_t tile0, const uint64_t *fix_off, const uint32_t *fix_cols) {
        const SymLaunch r = sym_prepare(L, n, det_dst, det_stride, obs_dst, obs_stride, tile0, 0, sym_scan && !fix_off);
        if (r.scan_ok) {
            check(launch_sym_scan(r.P, (uint32_t)r.tiles, r.lds_scan, stream, r.fast), "symptom scan launch");
        } else if (fix_off) {
            check(launch_sym_fixed(r.P, fix_off, fix_cols, (uint32_t)r.tiles, r.lds_queue, stream), "symptom fixed launch");
        } else {
            check(launch_sym_sample(r.P, (uint32_t)r.tiles, r.lds_queue, stream), "symptom sample launch");
        }
        return r.tiles;
    }

    void sample_symptom(const Layout &L, uint64_t shots, bool device_out, uint8_t *dets, size_t dets_stride,
                        uint8_t *obs, size_t obs_stride, const uint64_t *fault_offsets, const uint32_t *faults,
                        char base_char, bool newline, uint8_t *errs = nullptr, size_t errs_stride = 0) {
        const uint32_t O = prog.num_observables;
        // Fired-mechanism rows (detector error models, sampled mode): 32-bit words per row on the device.
        const bool want_err = errs && dem_mode && !faults;
        const uint64_t E = dem_col.size();
        const size_t ew = (E + 31) / 32, erow = L.packed ? (E + 7) / 8 : E;
        const size_t pdet = (L.ncols + 7) / 8, pobs = (O + 7) / 8;
        const bool want_obs = L.separate && obs && O;
        const bool piped = !device_out;
        uint64_t batch = shots;
        if (device_out && !L.packed) {
            // Batches are whole tiles of the launch's tile size, so the split never changes the tiles'
            // RNG streams.
            static uint8_t dummy;
            const uint64_t S = sym_prepare(L, 1, &dummy, pdet, want_obs ? &dummy : nullptr, pobs, 0).P.shots_per_tile;
            batch = std::max<uint64_t>(1024, HOST_BATCH_BYTES / std::max<size_t>(pdet + pobs + L.ncols + O, 1));
            batch = std::max<uint64_t>(S, batch / S * S);
            batch = arena_batch(batch, pdet + (want_obs ? pobs : 0), S);
        }
        if (want_err) {
            // Error rows take 32-bit words per mechanism block; batches stay whole tiles.
            static uint8_t dummy;
            const uint64_t S = sym_prepare(L, 1, &dummy, pdet, want_obs ? &dummy : nullptr, pobs, 0).P.shots_per_tile;
            const size_t per = 4 * ew + (L.packed ? 0 : E) + pdet + pobs;
            batch = std::min<uint64_t>(batch, std::max<uint64_t>(S, HOST_BATCH_BYTES / std::max<size_t>(per, 1) / S * S));
        }
        PipeCall pipe_call{this, piped};
        if (piped) {
            // Host rows travel through the pipeline: batches of about PIPE_BATCH_BYTES of host rows, in
            // whole tiles so that the split never changes the tiles' RNG streams.
            static uint8_t dummy;
            const uint64_t S = sym_prepare(L, 1, &dummy, pdet, want_obs ? &dummy : nullptr, pobs, 0).P.shots_per_tile;
            const size_t host_row = (L.packed ? pdet : L.ncols) + (want_obs ? (L.packed ? pobs : O) : 0) +
                                    (want_err ? erow : 0);
            batch = std::max<uint64_t>(S, PIPE_BATCH_BYTES / std::max<size_t>(host_row, 1) / S * S);
            batch = std::min(batch, shots);
            const size_t dev_row = std::max<size_t>(pdet, 1) + (want_obs ? std::max<size_t>(pobs, 1) : 0) +
                                   (L.packed ? 0 : std::max<size_t>(L.ncols + O, 1)) +
                                   (want_err ? 4 * ew + (L.packed ? 0 : E) : 0);
            pipe.begin(stream, batch * dev_row + 6 * 256, std::max<size_t>(batch * host_row, 256));
        }
        const uint64_t call_tile0 = tiles_used;
        uint64_t tiles_total = 0;
        for (uint64_t s0 = 0; s0 < shots; s0 += batch) {
            const uint64_t n = std::min(batch, shots - s0);
            // Packed destination rows for this batch.
            uint8_t *pd, *po = nullptr;
            size_t pd_stride, po_stride = 0;
            std::unique_ptr<SlotScope> slot;
            if (piped) slot.reset(new SlotScope(this));
            if (L.packed && device_out) {
                pd = dets + s0 * dets_stride;
                pd_stride = dets_stride;
                if (want_obs) {
                    po = obs + s0 * obs_stride;
                    po_stride = obs_stride;
                }
            } else {
                staging_reset();
                pd = staging(d_out_det, n * std::max<size_t>(pdet, 1));
                pd_stride = pdet;
                if (want_obs) {
                    po = staging(d_out_obs, n * std::max<size_t>(pobs, 1));
                    po_stride = pobs;
                }
            }
            const uint64_t *fo = nullptr;
            const uint32_t *fc = nullptr;
            if (faults) stage_symptom_faults(s0, n, fault_offsets, faults, &fo, &fc);
            uint8_t *err_dev = nullptr;
            if (want_err) {
                if (piped) {
                    err_dev = staging(d_err, n * ew * 4);
                } else {
                    d_err.ensure(n * ew * 4);
                    err_dev = d_err.as<uint8_t>();
                }
                check(hipMemsetAsync(err_dev, 0, n * ew * 4, stream), "clear error rows");
                err_words_cur = reinterpret_cast<uint32_t *>(err_dev);
                err_stride_cur = ew;
            }
            struct ClearErr {
                mfs_sampler *s;
                ~ClearErr() {
                    s->err_words_cur = nullptr;
                    s->err_stride_cur = 0;
                }
            } clear_err{this};
            tiles_total += sym_launch(L, n, pd, pd_stride, po, po_stride, call_tile0 + tiles_total, fo, fc);
            std::vector<HostPipe::Part> parts;
            if (want_err && E && piped) {
                // Little-endian 32-bit words hold the bits in the packed byte order.
                if (L.packed) {
                    parts.push_back({err_dev, ew * 4, erow, errs + s0 * errs_stride, errs_stride, -1L});
                } else {
                    uint8_t *eu = staging(d_err_tmp, n * E);
                    check(launch_unpack_rows(err_dev, ew * 4, eu, E, n, (uint32_t)E, 0, stream), "unpack error rows");
                    parts.push_back({eu, E, E, errs + s0 * errs_stride, errs_stride, -1L});
                }
            } else if (want_err && E) {
                // Device rows: copied (packed) or expanded to bytes in place.
                if (L.packed) {
                    check(copy_rows_async(errs + s0 * errs_stride, errs_stride, err_dev, ew * 4, erow, n,
                                           hipMemcpyDeviceToDevice, stream),
                          "copy error rows");
                } else {
                    check(launch_unpack_rows(err_dev, ew * 4, errs + s0 * errs_stride, errs_stride, n, (uint32_t)E, 0,
                                             stream), "unpack error rows");
                }
            }
            if (!L.packed) {
                // Expand bits to bytes (bool) or characters.
                uint8_t *ud = device_out ? dets + s0 * dets_stride : nullptr;
                uint8_t *uo = device_out && want_obs ? obs + s0 * obs_stride : nullptr;
                size_t ud_stride = dets_stride, uo_stride = obs_stride;
                if (!device_out) {
                    ud = staging(d_tmp, n * std::max<size_t>(L.ncols + O, 1));
                    ud_stride = L.ncols;
                    uo = ud + n * L.ncols;
                    uo_stride = O;
                }
                check(launch_unpack_rows(pd, pd_stride, ud, ud_stride, n, L.ncols, (uint8_t)base_char, stream), "unpack");
                if (want_obs) check(launch_unpack_rows(po, po_stride, uo, uo_stride, n, O, (uint8_t)base_char, stream), "unpack");
                pd = ud;
                pd_stride = ud_stride;
                po = uo;
                po_stride = uo_stride;
            }
            if (piped) {
                const size_t drow = L.packed ? pdet : L.ncols, orow = L.packed ? pobs : O;
                if (dets && (drow || newline)) {
                    parts.push_back({pd, pd_stride, drow, dets + s0 * dets_stride, dets_stride,
                                     newline ? (long)L.ncols : -1L});
                }
                if (want_obs && orow) parts.push_back({po, po_stride, orow, obs + s0 * obs_stride, obs_stride, -1L});
                pipe.submit(parts, n);
            }
        }
        if (piped) pipe.finish();
        check(hipStreamSynchronize(stream), "sample");
        tiles_used += tiles_total;
    }

    // Maps fixed faults to symptom columns (per shot, the basis columns of each fault's code).
    void stage_symptom_faults(uint64_t s0, uint64_t n, const uint64_t *off, const uint32_t *faults,
                              const uint64_t **fo, const uint32_t **fc) {
        std::vector<uint64_t> offs(n + 1, 0);
        std::vector<uint32_t> cols;
        for (uint64_t s = 0; s < n; s++) {
            for (uint64_t i = off[s0 + s]; i < off[s0 + s + 1]; i++) {
                if (dem_mode) {
                    // (mechanism index, 0, 1): the mechanism fires in this shot.
                    const uint32_t e = faults[3 * i];
                    if (e >= dem_col.size() || faults[3 * i + 1] != 0 || faults[3 * i + 2] > 1) {
                        throw std::out_of_range("error mechanism out of range");
                    }
                    if (faults[3 * i + 2]) cols.push_back(dem_col[e]);
                    continue;
                }
                uint32_t op, site;
                resolve_fault(prog, faults[3 * i], faults[3 * i + 1], &op, &site);
                const uint32_t c = plan.op_class[op];
                if (c == 0xFFFFFFFFu) {
                    throw std::invalid_argument("the symptom engine has no column for a fault at a zero-probability site");
                }
                const SymClass &cl = plan.classes[c];
                const uint32_t code = faults[3 * i + 2];
                if ((prog.ops[op].code & 0xFF) == OP_CORR) {
                    if (code & 1) cols.push_back(cl.col0 + plan.op_link[op]);
                    continue;
                }
                const uint32_t cs = plan.op_site0[op] + site;
                for (uint32_t j = 0; j < cl.basis; j++) {
                    if (code & (1u << j)) cols.push_back(cl.col0 + cs * cl.basis + j);
                }
            }
            offs[s + 1] = cols.size();
        }
        d_fix_off.ensure(offs.size() * 8);
        d_fix_cols.ensure(std::max<size_t>(cols.size(), 1) * 4);
        check(hipMemcpyAsync(d_fix_off.p, offs.data(), offs.size() * 8, hipMemcpyHostToDevice, stream), "fault offsets");
        if (!cols.empty())
            check(hipMemcpyAsync(d_fix_cols.p, cols.data(), cols.size() * 4, hipMemcpyHostToDevice, stream), "fault cols");
        *fo = d_fix_off.as<uint64_t>();
        *fc = d_fix_cols.as<uint32_t>();
    }

    void stage_faults(uint64_t s0, uint64_t n, uint64_t tiles, const uint64_t *off, const uint32_t *faults,
                      LaunchParams &P) {
        std::vector<DevFault> fs;
        std::vector<uint32_t> fo(tiles + 1, 0);
        const uint32_t W32 = W();
        for (uint64_t t = 0; t < tiles; t++) {
            fo[t] = (uint32_t)fs.size();
            const size_t start = fs.size();
            for (uint64_t s = s0 + t * W32; s < std::min(s0 + n, s0 + (t + 1) * W32); s++) {
                for (uint64_t i = off[s]; i < off[s + 1]; i++) {
                    uint32_t op, site;
                    resolve_fault(prog, faults[3 * i], faults[3 * i + 1], &op, &site);
                    fs.push_back(DevFault{op, site, (uint32_t)(s - s0 - t * W32), faults[3 * i + 2]});
                }
            }
            std::stable_sort(fs.begin() + (long)start, fs.end(),
                             [](const DevFault &a, const DevFault &b) { return a.op < b.op; });
        }
        fo[tiles] = (uint32_t)fs.size();
        d_faults.ensure(std::max<size_t>(fs.size(), 1) * sizeof(DevFault));
        d_fault_off.ensure(fo.size() * 4);
        if (!fs.empty()) check(hipMemcpy(d_faults.p, fs.data(), fs.size() * sizeof(DevFault), hipMemcpyHostToDevice), "faults");
        check(hipMemcpy(d_fault_off.p, fo.data(), fo.size() * 4, hipMemcpyHostToDevice), "fault offsets");
        P.faults = d_faults.as<DevFault>();
        P.fault_off = d_fault_off.as<uint32_t>();
    }
};

// ---------------------------------------------------------------------------------------------
// Result-format writers (src/stim/io/measure_record_writer.cc:23-211, measure_record_writer.h:111-167).

