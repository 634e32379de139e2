// C-ABI entry points of the stim_amd sampler: compile, sample, sample_write, fixed-fault sampling.
//
// Mirrors the reference's CompiledDetectorSampler (src/stim/py/compiled_detector_sampler.pybind.{h,cc})
// and its drivers (src/stim/simulators/frame_simulator_util.inl): the circuit is compiled once into a
// device program; each call simulates batches of shot tiles on the GPU, writing shot-major rows
// directly (no separate transpose pass), and copies them to the caller.
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
#include <vector>

#include "../../include/stim_amd.h"
#include "circuit.h"
#include "compiler.h"
#include "program.h"
#include "symptom.h"

namespace sa {
hipError_t launch_frame_kernel(int T, const DevProgram &prog, const LaunchParams &P, uint32_t num_tiles, size_t lds,
                               hipStream_t stream);
hipError_t launch_pack_rows(const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride, uint64_t rows,
                            uint32_t ncols, hipStream_t stream);
hipError_t launch_sym_sample(const SymDevParams &P, uint32_t tiles, size_t lds, hipStream_t stream);
hipError_t launch_sym_fixed(const SymDevParams &P, const uint64_t *off, const uint32_t *cols, uint32_t tiles,
                            size_t lds, hipStream_t stream);
hipError_t launch_sym_scan(const SymDevParams &P, uint32_t tiles, size_t lds, hipStream_t stream);
hipError_t launch_sym_batch(SymDevParams *jobs, const uint64_t *prefix, uint32_t njobs, uint32_t *counter,
                            uint32_t grid, size_t lds, hipStream_t stream);
int sym_batch_blocks_per_cu(size_t lds);
hipError_t launch_sym_count(const uint8_t *det, uint32_t det_row, const uint8_t *obs, uint32_t obs_row, uint64_t rows,
                            uint32_t *counts, hipStream_t stream);
hipError_t launch_sym_fill(const uint8_t *det, uint32_t det_row, const uint8_t *obs, uint32_t obs_row, uint64_t rows,
                           uint32_t D, const uint32_t *offs, uint32_t *ent, hipStream_t stream);
hipError_t launch_unpack_rows(const uint8_t *src, uint64_t src_stride, uint8_t *dst, uint64_t dst_stride, uint64_t rows,
                              uint32_t ncols, uint32_t base_char, hipStream_t stream);
}  // namespace sa

using namespace sa;

namespace {

thread_local std::string g_last_error;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char *what) {
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

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

constexpr size_t MAX_LDS = 160 * 1024;
constexpr size_t HOST_BATCH_BYTES = (size_t)1 << 30;

}  // namespace

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
    DevBuf d_classes, d_sotab, d_recs, d_ovf, d_sym_tmp, d_fix_off, d_fix_cols;
    int sym_tile_kb = 32;        // target LDS image size per symptom tile
    int sym_scan = 1;            // scan kernel (0: per-lane-walk queue kernel)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ~mfs_sampler() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
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
                bool newline = false) {
        Layout L = layout(flags);
        if (L.separate && (L.prepend || L.append)) {
            throw std::invalid_argument(
                "Can't specify separate_observables=True with append_observables=True or prepend_observables=True");
        }
        const bool device_out = flags & MFS_DEVICE_OUTPUT;
        if (shots == 0) return;
        if (use_symptoms(faults != nullptr)) {
            sample_symptom(L, shots, device_out, dets, dets_stride, obs, obs_stride, fault_offsets, faults, base_char,
                           newline);
            return;
        }
        const uint64_t W64 = W();
        const size_t row_bytes = L.det_row + (L.separate ? L.obs_row : 0) + (L.via_unpacked ? L.ncols : 0);
        uint64_t batch = device_out && !L.via_unpacked ? shots
                                                        : std::max<uint64_t>(W64, (HOST_BATCH_BYTES / std::max<size_t>(row_bytes, 1)) / W64 * W64);
        batch = std::min(batch, shots);
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
            uint8_t *det_dst, *obs_dst;
            size_t det_stride_dev, obs_stride_dev;
            if (L.via_unpacked) {
                d_tmp.ensure(n * L.ncols);
                Layout U = L;
                U.packed = false;
                U.via_unpacked = false;
                outputs(U, d_tmp.as<uint8_t>(), L.ncols, nullptr, 0, P);
                if (device_out) {
                    det_dst = dets + s0 * dets_stride;
                    det_stride_dev = dets_stride;
                } else {
                    d_out_det.ensure(n * L.det_row);
                    det_dst = d_out_det.as<uint8_t>();
                    det_stride_dev = L.det_row;
                }
                launch(P, tiles);
                check(launch_pack_rows(d_tmp.as<uint8_t>(), L.ncols, det_dst, det_stride_dev, n, L.ncols, stream),
                      "pack rows");
                obs_dst = nullptr;
                obs_stride_dev = 0;
            } else {
                if (device_out) {
                    det_dst = dets + s0 * dets_stride;
                    det_stride_dev = dets_stride;
                    obs_dst = (L.separate && obs) ? obs + s0 * obs_stride : nullptr;
                    obs_stride_dev = obs_stride;
                } else {
                    d_out_det.ensure(n * std::max<size_t>(L.det_row, 1));
                    det_dst = d_out_det.as<uint8_t>();
                    det_stride_dev = L.det_row;
                    d_out_obs.ensure(n * std::max<size_t>(L.obs_row, 1));
                    obs_dst = L.separate ? d_out_obs.as<uint8_t>() : nullptr;
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
            if (!device_out) {
                if (L.det_row && dets) {
                    check(hipMemcpy2DAsync(dets + s0 * dets_stride, dets_stride, det_dst, det_stride_dev, L.det_row, n,
                                           hipMemcpyDeviceToHost, stream),
                          "copy dets");
                }
                if (L.separate && obs && L.obs_row && obs_dst) {
                    check(hipMemcpy2DAsync(obs + s0 * obs_stride, obs_stride, obs_dst, obs_stride_dev, L.obs_row, n,
                                           hipMemcpyDeviceToHost, stream),
                          "copy obs");
                }
                check(hipStreamSynchronize(stream), "sample");
                if (newline) {
                    for (uint64_t i = 0; i < n; i++) dets[(s0 + i) * dets_stride + L.ncols] = '\n';
                }
            }
        }
        if (device_out) check(hipStreamSynchronize(stream), "sample");
        tiles_used += (shots + W64 - 1) / W64;
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
        sym_state = 1;
        return true;
    }

    bool use_symptoms(bool fixed) {
        if (engine_mode == 1) return false;
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
        const bool ok = ensure_symptoms();
        if (!ok && engine_mode == 2) throw std::invalid_argument("symptom engine unavailable: " + sym_why);
        return ok;
    }

    // Shots per symptom tile: LDS image of about 32 KB.
    uint32_t sym_tile_shots(size_t row_bytes, int kb) const {
        uint32_t S = 1;
        while (S < 1024 && (size_t)(2 * S) * row_bytes <= (size_t)kb * 1024) S *= 2;
        return S;
    }

    SymDevParams sym_params(uint32_t S) {
        SymDevParams P{};
        P.classes = d_classes.as<SymClass>();
        P.otab = d_sotab.as<uint32_t>();
        P.recs = d_recs.as<uint2>();
        P.ovf = d_ovf.as<uint32_t>();
        P.nclasses = (uint32_t)plan.classes.size();
        P.otab_words = (uint32_t)plan.otab.size();
        P.D = prog.num_detectors;
        P.O = prog.num_observables;
        P.shots_per_tile = S;
        uint32_t lg = 0;
        while ((1u << lg) < S) lg++;
        P.tile_log2 = lg;
        P.total_sites = plan.total_sites;
        P.seed = splitmix64(key ^ 0x5A5A5A5Aull);
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
    };

    SymLaunch sym_prepare(const Layout &L, uint64_t n, uint8_t *det_dst, size_t det_stride, uint8_t *obs_dst,
                          size_t obs_stride, uint64_t tile0, int tile_kb = 0) {
        const uint32_t D = prog.num_detectors, O = prog.num_observables;
        const size_t obs_bytes = (L.separate && obs_dst) ? (O + 7) / 8 : 0;
        const size_t det_bytes = (L.ncols + 7) / 8;
        const uint32_t S = sym_tile_shots(std::max<size_t>(det_bytes + obs_bytes, 1), tile_kb ? tile_kb : sym_tile_kb);
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
        r.scan_ok = (plan.total_sites << P.tile_log2) < (1ull << 31);
        return r;
    }

    // Launches the symptom engine over n shots starting at RNG tile tile0; returns the tiles used.
    uint64_t sym_launch(const Layout &L, uint64_t n, uint8_t *det_dst, size_t det_stride, uint8_t *obs_dst,
                        size_t obs_stride, uint64_t tile0, const uint64_t *fix_off, const uint32_t *fix_cols) {
        const SymLaunch r = sym_prepare(L, n, det_dst, det_stride, obs_dst, obs_stride, tile0);
        if (sym_scan && !fix_off && r.scan_ok) {
            check(launch_sym_scan(r.P, (uint32_t)r.tiles, r.lds_scan, stream), "symptom scan launch");
        } else if (fix_off) {
            check(launch_sym_fixed(r.P, fix_off, fix_cols, (uint32_t)r.tiles, r.lds_queue, stream), "symptom fixed launch");
        } else {
            check(launch_sym_sample(r.P, (uint32_t)r.tiles, r.lds_queue, stream), "symptom sample launch");
        }
        return r.tiles;
    }

    void sample_symptom(const Layout &L, uint64_t shots, bool device_out, uint8_t *dets, size_t dets_stride,
                        uint8_t *obs, size_t obs_stride, const uint64_t *fault_offsets, const uint32_t *faults,
                        char base_char, bool newline) {
        const uint32_t O = prog.num_observables;
        const size_t pdet = (L.ncols + 7) / 8, pobs = (O + 7) / 8;
        const bool want_obs = L.separate && obs && O;
        const uint64_t batch = device_out && L.packed
                                   ? shots
                                   : std::max<uint64_t>(1024, HOST_BATCH_BYTES / std::max<size_t>(pdet + pobs + L.ncols + O, 1));
        const uint64_t call_tile0 = tiles_used;
        uint64_t tiles_total = 0;
        for (uint64_t s0 = 0; s0 < shots; s0 += batch) {
            const uint64_t n = std::min(batch, shots - s0);
            // Packed destination rows for this batch.
            uint8_t *pd, *po = nullptr;
            size_t pd_stride, po_stride = 0;
            if (L.packed && device_out) {
                pd = dets + s0 * dets_stride;
                pd_stride = dets_stride;
                if (want_obs) {
                    po = obs + s0 * obs_stride;
                    po_stride = obs_stride;
                }
            } else {
                d_out_det.ensure(n * std::max<size_t>(pdet, 1));
                pd = d_out_det.as<uint8_t>();
                pd_stride = pdet;
                if (want_obs) {
                    d_out_obs.ensure(n * std::max<size_t>(pobs, 1));
                    po = d_out_obs.as<uint8_t>();
                    po_stride = pobs;
                }
            }
            const uint64_t *fo = nullptr;
            const uint32_t *fc = nullptr;
            if (faults) stage_symptom_faults(s0, n, fault_offsets, faults, &fo, &fc);
            tiles_total += sym_launch(L, n, pd, pd_stride, po, po_stride, call_tile0 + tiles_total, fo, fc);
            if (!L.packed) {
                // Expand bits to bytes (bool) or characters.
                uint8_t *ud = device_out ? dets + s0 * dets_stride : nullptr;
                uint8_t *uo = device_out && want_obs ? obs + s0 * obs_stride : nullptr;
                size_t ud_stride = dets_stride, uo_stride = obs_stride;
                if (!device_out) {
                    d_tmp.ensure(n * std::max<size_t>(L.ncols + O, 1));
                    ud = d_tmp.as<uint8_t>();
                    ud_stride = L.ncols;
                    uo = d_tmp.as<uint8_t>() + n * L.ncols;
                    uo_stride = O;
                }
                check(launch_unpack_rows(pd, pd_stride, ud, ud_stride, n, L.ncols, (uint8_t)base_char, stream), "unpack");
                if (want_obs) check(launch_unpack_rows(po, po_stride, uo, uo_stride, n, O, (uint8_t)base_char, stream), "unpack");
                pd = ud;
                pd_stride = ud_stride;
                po = uo;
                po_stride = uo_stride;
            }
            if (!device_out) {
                const size_t drow = L.packed ? pdet : L.ncols, orow = L.packed ? pobs : O;
                if (drow && dets)
                    check(hipMemcpy2DAsync(dets + s0 * dets_stride, dets_stride, pd, pd_stride, drow, n,
                                           hipMemcpyDeviceToHost, stream), "copy dets");
                if (want_obs && orow)
                    check(hipMemcpy2DAsync(obs + s0 * obs_stride, obs_stride, po, po_stride, orow, n,
                                           hipMemcpyDeviceToHost, stream), "copy obs");
                check(hipStreamSynchronize(stream), "sample");
                if (newline) {
                    for (uint64_t i = 0; i < n; i++) dets[(s0 + i) * dets_stride + L.ncols] = '\n';
                }
            }
        }
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
    std::string line;
    for (uint64_t s = 0; s < nshots; s++) {
        const uint8_t *r = rows + s * stride;
        auto bit = [&](uint32_t m) { return (r[m >> 3] >> (m & 7)) & 1; };
        switch (fmt) {
            case Fmt::B8:
                fwrite(r, 1, (ncols + 7) / 8, out);
                break;
            case Fmt::F01:
                line.assign(ncols + 1, '\n');
                for (uint32_t m = 0; m < ncols; m++) line[m] = (char)('0' + bit(m));
                fwrite(line.data(), 1, line.size(), out);
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

int fail(const std::exception &e, int kind) {
    g_last_error = e.what();
    return kind;
}

template <class F>
int guarded(F f) {
    try {
        f();
        return MFS_OK;
    } catch (const std::out_of_range &e) {
        return fail(e, MFS_ERR_OUT_OF_RANGE);
    } catch (const std::invalid_argument &e) {
        return fail(e, MFS_ERR_INVALID_ARGUMENT);
    } catch (const std::exception &e) {
        return fail(e, MFS_ERR_RUNTIME);
    }
}

}  // namespace

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
        std::vector<uint8_t> dbuf, obuf;
        char p1 = 'D', p2 = 'L';
        uint32_t transition = D;
        if (prepend_observables) {
            p1 = 'L';
            p2 = 'D';
            transition = O;
        }
        for (uint64_t s0 = 0; s0 < shots; s0 += chunk) {
            const uint64_t n = std::min(chunk, shots - s0);
            dbuf.assign(n * std::max<size_t>(drow, 1), 0);
            obuf.assign(n * std::max<size_t>(orow, 1), 0);
            s->sample(n, flags, dbuf.data(), std::max<size_t>(drow, 1), obuf.data(), std::max<size_t>(orow, 1));
            if (fmt == Fmt::PTB64) write_ptb64(out.get(), dbuf.data(), std::max<size_t>(drow, 1), n, ncols);
            else write_rows(out.get(), fmt, dbuf.data(), std::max<size_t>(drow, 1), n, ncols, p1, p2, transition);
            if (has_obs) {
                if (ofmt == Fmt::PTB64) write_ptb64(oout.get(), obuf.data(), std::max<size_t>(orow, 1), n, O);
                else write_rows(oout.get(), ofmt, obuf.data(), std::max<size_t>(orow, 1), n, O, 'L', 'L', O);
            }
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
        const bool sym = s->use_symptoms(false);
        auto one = [&]() {
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
        }
        out_ms[0] = iters ? kernel_total / iters : 0;
        out_ms[1] = out_ms[0];
    });
}

int mfs_set_tile_words(mfs_sampler *s, int words) {
    int rc = guarded([&] {
        if (words != 0 && words != 1 && words != 2 && words != 4) throw std::invalid_argument("tile words must be 0, 1, 2 or 4");
        s->pick_tile_words(words);
    });
    if (rc) return -rc;
    return s->T;
}

int mfs_set_engine(mfs_sampler *s, int engine) {
    return guarded([&] {
        if (engine < 0 || engine > 2) throw std::invalid_argument("engine must be 0 (auto), 1 (frame) or 2 (symptom)");
        s->engine_mode = engine;
        if (engine == 2 && !s->use_symptoms(false)) throw std::invalid_argument("symptom engine unavailable");
    });
}

int mfs_engine_info(mfs_sampler *s, uint64_t out[6]) {
    return guarded([&] {
        const bool sym = s->use_symptoms(false);
        out[0] = sym ? 2 : 1;
        out[1] = s->plan.classes.size();
        out[2] = s->plan.columns.size();
        out[3] = s->sym_entries;
        out[4] = (uint64_t)(s->sym_build_ms * 1000.0);
        out[5] = s->plan.total_sites;
    });
}

int mfs_set_symptom_tile_kb(mfs_sampler *s, int kb) {
    return guarded([&] {
        if (kb == 0 || kb < -128 || kb > 128) throw std::invalid_argument("symptom tile size must be 1..128 KB");
        // A negative size selects the per-lane-walk queue kernel instead of the scan kernel.
        s->sym_scan = kb > 0;
        s->sym_tile_kb = kb > 0 ? kb : -kb;
    });
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

// ------------------------------------------------------------------------------------------------
// Batches: many sampling jobs (each a sampler, a shot count and device output rows) run by one
// persistent kernel. Jobs of one batch may come from different circuits. Each job draws the same
// random stream as mfs_sample on its sampler would at the same point of the sampler's stream, with the
// same tile size, so a batch run equals the separate launches bit for bit.

struct mfs_batch {
    struct Job {
        mfs_sampler *s;
        uint64_t shots;
        uint8_t *dets;
        size_t dets_stride;
        uint8_t *obs;
        size_t obs_stride;
    };
    std::vector<Job> jobs;
    int tile_kb = 0;                 // LDS output image per workgroup (0: each sampler's own setting)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf d_jobs, d_prefix, d_counter;
    // Prepared state.
    bool prepared = false;
    std::vector<uint32_t> sym_jobs;      // job indices run by the persistent kernel, in launch order
    std::vector<uint32_t> frame_jobs;    // job indices run as their own launches
    std::vector<std::pair<mfs_sampler *, uint64_t>> advance;  // RNG tiles each sampler uses per run
    std::vector<std::pair<mfs_sampler *, uint64_t>> expect;   // sampler tile counters the params assume
    uint64_t total_tiles = 0;
    size_t lds = 0;
    uint32_t grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~mfs_batch() {
        drop_graph();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    void invalidate() {
        prepared = false;
        drop_graph();
    }

    void prepare() {
        invalidate();
        sym_jobs.clear();
        frame_jobs.clear();
        advance.clear();
        expect.clear();
        // Per sampler: the tile counter at the start of the run and the tiles its jobs take.
        std::vector<std::pair<mfs_sampler *, uint64_t>> next_tile;
        auto slot = [&](mfs_sampler *s) -> uint64_t & {
            for (auto &kv : next_tile)
                if (kv.first == s) return kv.second;
            next_tile.push_back({s, s->tiles_used});
            expect.push_back({s, s->tiles_used});
            return next_tile.back().second;
        };
        struct Item {
            SymDevParams P;
            uint64_t tiles;
            double cost;
            uint32_t job;
        };
        std::vector<Item> items;
        size_t max_lds = 0;
        for (uint32_t j = 0; j < jobs.size(); j++) {
            Job &jb = jobs[j];
            mfs_sampler *s = jb.s;
            const bool sym = s->use_symptoms(false);
            if (!sym) {
                frame_jobs.push_back(j);
                continue;
            }
            check(hipStreamSynchronize(s->stream), "symptom tables");
            const uint32_t flags = MFS_BIT_PACKED | MFS_DEVICE_OUTPUT | (jb.obs ? MFS_SEPARATE_OBSERVABLES : 0);
            const auto L = s->layout(flags);
            auto r = s->sym_prepare(L, jb.shots, jb.dets, jb.dets_stride, jb.obs, jb.obs_stride, 0, tile_kb);
            if (!r.scan_ok || !s->sym_scan) {
                frame_jobs.push_back(j);
                continue;
            }
            uint64_t &t0 = slot(s);
            r.P.tile0 = t0;
            t0 += r.tiles;
            double hits = 0;
            for (const SymClass &cl : s->plan.classes) hits += (double)cl.nsites * cl.p;
            const double S = r.P.shots_per_tile;
            // Per-tile cost estimate: hits applied plus output bytes written, plus a fixed overhead.
            const double cost = S * (hits * 16.0 + r.P.det_row_bytes + r.P.obs_row_bytes) + 4096.0;
            items.push_back({r.P, r.tiles, cost, j});
            max_lds = std::max(max_lds, r.lds_scan);
        }
        for (auto &kv : next_tile) {
            uint64_t start = 0;
            for (auto &e : expect)
                if (e.first == kv.first) start = e.second;
            advance.push_back({kv.first, kv.second - start});
        }
        // Heaviest tiles first, so the tail of the run is made of light tiles.
        std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
        std::vector<SymDevParams> params;
        std::vector<uint64_t> prefix{0};
        for (const Item &it : items) {
            params.push_back(it.P);
            prefix.push_back(prefix.back() + it.tiles);
            sym_jobs.push_back(it.job);
        }
        total_tiles = prefix.back();
        if (total_tiles >= (1ull << 32)) throw std::invalid_argument("too many tiles in one batch");
        if (max_lds > MAX_LDS) throw std::invalid_argument("a batch job's tile does not fit in LDS");
        lds = std::max<size_t>(max_lds, 64);
        if (!stream) {
            check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
            check(hipEventCreate(&ev0), "event");
            check(hipEventCreate(&ev1), "event");
        }
        if (!params.empty()) {
            d_jobs.ensure(params.size() * sizeof(SymDevParams));
            d_prefix.ensure(prefix.size() * sizeof(uint64_t));
            d_counter.ensure(4);
            check(hipMemcpyAsync(d_jobs.p, params.data(), params.size() * sizeof(SymDevParams), hipMemcpyHostToDevice,
                                 stream), "upload batch");
            check(hipMemcpyAsync(d_prefix.p, prefix.data(), prefix.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                                 stream), "upload batch");
            check(hipStreamSynchronize(stream), "upload batch");
            int dev = 0, cus = 0;
            check(hipGetDevice(&dev), "device");
            check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device");
            const uint64_t resident = (uint64_t)std::max(cus, 1) * (uint64_t)sym_batch_blocks_per_cu(lds);
            grid = (uint32_t)std::min<uint64_t>(resident, total_tiles);
        } else {
            grid = 0;
        }
        prepared = true;
    }

    // Re-prepares when a sampler's tile counter moved outside this batch (a plain call in between).
    void sync_counters() {
        if (!prepared) {
            prepare();
            return;
        }
        for (auto &e : expect) {
            if (e.first->tiles_used != e.second) {
                prepare();
                return;
            }
        }
    }

    void enqueue_sym() {
        if (!grid) return;
        check(launch_sym_batch(d_jobs.as<SymDevParams>(), d_prefix.as<uint64_t>(), (uint32_t)sym_jobs.size(),
                               d_counter.as<uint32_t>(), grid, lds, stream), "batch launch");
    }

    // One run: the persistent kernel (plain launches or a graph replay), then the frame-engine jobs.
    void run_once(bool use_graph) {
        sync_counters();
        if (grid) {
            if (use_graph) {
                if (!exec) {
                    check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "graph capture");
                    enqueue_sym();
                    check(hipStreamEndCapture(stream, &graph), "graph capture");
                    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
                }
                check(hipGraphLaunch(exec, stream), "graph launch");
            } else {
                enqueue_sym();
            }
        }
        for (auto &a : advance) a.first->tiles_used += a.second;
        for (auto &e : expect) {
            for (auto &a : advance)
                if (a.first == e.first) e.second += a.second;
        }
        for (uint32_t j : frame_jobs) {
            Job &jb = jobs[j];
            const uint32_t flags = MFS_BIT_PACKED | MFS_DEVICE_OUTPUT | (jb.obs ? MFS_SEPARATE_OBSERVABLES : 0);
            jb.s->sample(jb.shots, flags, jb.dets, jb.dets_stride, jb.obs, jb.obs_stride);
            // A sampler with jobs on both paths: its device tile bases are now stale.
            for (auto &e : expect)
                if (e.first == jb.s) invalidate();
        }
    }
};

extern "C" {

int mfs_batch_create(mfs_batch **out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = new mfs_batch();
    });
}

void mfs_batch_destroy(mfs_batch *b) { delete b; }

int mfs_batch_add(mfs_batch *b, mfs_sampler *s, uint64_t shots, uint8_t *dets, size_t dets_stride, uint8_t *obs,
                  size_t obs_stride) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("sampler is null");
        if (!dets && shots) throw std::invalid_argument("dets is null");
        const size_t drow = (s->prog.num_detectors + 7) / 8, orow = (s->prog.num_observables + 7) / 8;
        if (dets_stride < drow) throw std::invalid_argument("dets_stride is smaller than a packed detector row");
        if (obs && obs_stride < orow) throw std::invalid_argument("obs_stride is smaller than a packed observable row");
        b->jobs.push_back({s, shots, dets, dets_stride, obs, obs_stride});
        b->invalidate();
    });
}

int mfs_batch_set_tile_kb(mfs_batch *b, int kb) {
    return guarded([&] {
        if (kb < 0 || kb > 128) throw std::invalid_argument("batch tile size must be 0..128 KB");
        b->tile_kb = kb;
        b->invalidate();
    });
}

int mfs_batch_run(mfs_batch *b, int use_graph) {
    return guarded([&] {
        b->run_once(use_graph != 0);
        check(hipStreamSynchronize(b->stream ? b->stream : nullptr), "batch run");
    });
}

int mfs_batch_time(mfs_batch *b, int use_graph, int warmup, int iters, double out_ms[2]) {
    return guarded([&] {
        if (iters < 1 || warmup < 0) throw std::invalid_argument("iters must be >= 1 and warmup >= 0");
        for (int i = 0; i < warmup; i++) b->run_once(use_graph != 0);
        b->sync_counters();
        check(hipStreamSynchronize(b->stream), "warmup");
        const auto t0 = std::chrono::steady_clock::now();
        check(hipEventRecord(b->ev0, b->stream), "event");
        for (int i = 0; i < iters; i++) b->run_once(use_graph != 0);
        check(hipEventRecord(b->ev1, b->stream), "event");
        check(hipEventSynchronize(b->ev1), "event");
        const auto t1 = std::chrono::steady_clock::now();
        float ms = 0;
        check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "elapsed");
        out_ms[0] = ms / iters;
        out_ms[1] = std::chrono::duration<double, std::milli>(t1 - t0).count() / iters;
    });
}

int mfs_batch_autotune(mfs_batch *b, const int *candidates, int n, int iters, int *best_kb, double *best_ms) {
    return guarded([&] {
        if (!candidates || n < 1) throw std::invalid_argument("no candidate tile sizes");
        double best = 1e300;
        int pick = candidates[0];
        for (int i = 0; i < n; i++) {
            if (candidates[i] < 1 || candidates[i] > 128) throw std::invalid_argument("candidate tile sizes must be 1..128 KB");
            b->tile_kb = candidates[i];
            b->invalidate();
            b->run_once(false);
            check(hipStreamSynchronize(b->stream), "autotune");
            const auto t0 = std::chrono::steady_clock::now();
            for (int k = 0; k < std::max(iters, 1); k++) b->run_once(false);
            check(hipStreamSynchronize(b->stream), "autotune");
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                              std::max(iters, 1);
            if (ms < best) {
                best = ms;
                pick = candidates[i];
            }
        }
        b->tile_kb = pick;
        b->invalidate();
        // Plain calls on these samplers use the chosen tiles too, so they stay comparable to the batch.
        for (auto &jb : b->jobs) jb.s->sym_tile_kb = pick;
        if (best_kb) *best_kb = pick;
        if (best_ms) *best_ms = best;
    });
}

int mfs_batch_info(mfs_batch *b, uint64_t out[6]) {
    return guarded([&] {
        b->sync_counters();
        out[0] = b->jobs.size();
        out[1] = b->sym_jobs.size();
        out[2] = b->frame_jobs.size();
        out[3] = b->total_tiles;
        out[4] = b->grid;
        out[5] = b->lds;
    });
}

}  // extern "C"
