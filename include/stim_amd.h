/* stim_amd: MI355X-native detection-event sampler, C ABI.
 *
 * Replaces, for the bulk Pauli-frame detection-event path, the reference's
 *   stim.Circuit.compile_detector_sampler(seed=...)
 *       src/stim/circuit/circuit.pybind.cc:827-830, src/stim/py/compiled_detector_sampler.pybind.cc:151-154
 *   CompiledDetectorSampler.sample(shots, prepend/append/separate_observables, bit_packed, dets_out, obs_out)
 *       src/stim/py/compiled_detector_sampler.pybind.cc:34-86, 216-322
 *   CompiledDetectorSampler.sample_write(shots, filepath, format, ..., obs_out_filepath, obs_out_format)
 *       src/stim/py/compiled_detector_sampler.pybind.cc:88-141, 357-415
 *   CompiledDetectorSampler.sample_bit_packed(...)   :324-355 (legacy alias)
 * Plain pointers and sizes only; every call returns 0 on success or a non-zero error kind, with the
 * message available from mfs_last_error() (thread-local). Error kinds map to the exceptions the
 * reference raises through pybind11: MFS_ERR_INVALID_ARGUMENT -> ValueError (std::invalid_argument),
 * MFS_ERR_OUT_OF_RANGE -> IndexError (std::out_of_range).
 *
 * A sampler handle is single-threaded (like the reference object); use one handle per thread.
 */
#ifndef STIM_AMD_H
#define STIM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfs_sampler mfs_sampler;

enum {
    MFS_OK = 0,
    MFS_ERR_INVALID_ARGUMENT = 1, /* ValueError */
    MFS_ERR_OUT_OF_RANGE = 2,     /* IndexError */
    MFS_ERR_RUNTIME = 3,          /* RuntimeError (HIP failures, no GPU, ...) */
};

/* Flags for mfs_sample / mfs_sample_with_faults. */
enum {
    MFS_BIT_PACKED = 1u << 0,        /* (S, ceil(n/8)) uint8, bit m of shot s at byte m/8 bit m%8; else (S, n) bytes 0/1 */
    MFS_PREPEND_OBSERVABLES = 1u << 1,
    MFS_APPEND_OBSERVABLES = 1u << 2,
    MFS_SEPARATE_OBSERVABLES = 1u << 3, /* write observables to obs (reference: separate_observables or obs_out) */
    MFS_DEVICE_OUTPUT = 1u << 4,     /* dets/obs are device pointers (results stay in HBM) */
};

/* Compiles .stim circuit text. seed is used when has_seed != 0 (reference: seed=None draws
 * entropy from the OS, probability_util.cc:53-64). */
int mfs_compile(const char *stim_text, size_t len, uint64_t seed, int has_seed, mfs_sampler **out);

void mfs_destroy(mfs_sampler *s);

/* out[0..5] = num_qubits, num_measurements, num_detectors, num_observables, max_lookback. */
int mfs_circuit_stats(const mfs_sampler *s, uint64_t out[5]);

/* Parses and compiles circuit text on the host only (no GPU needed) and reports
 * out[0..9] = num_qubits, num_measurements, num_detectors, num_observables, max_lookback,
 * device ops, barriers, noise ops, record ring rows, LDS bytes for a 64-shot tile.
 * Reference: Circuit.compute_stats / num_detectors etc. (src/stim/circuit/circuit_instruction.cc:44-100). */
int mfs_parse_stats(const char *stim_text, size_t len, uint64_t out[10]);

/* Compiles a measurement sampler: sampling returns every measurement result (record order) instead of
 * detection events, as the reference sample XOR the frame simulator's flips. The reference sample is
 * computed on the host (skip_reference_sample = 0, reference_sample_bits = NULL), taken as all zeros
 * (skip_reference_sample = 1), or given bit-packed little-endian in reference_sample_bits (at least
 * ceil(num_measurements / 8) bytes). mfs_sample / mfs_sample_with_faults / mfs_sample_write then take
 * ncols = num_measurements and no observables. Runs on the frame engine.
 * Reference: Circuit.compile_sampler -> CompiledMeasurementSampler (src/stim/circuit/circuit.pybind.cc:588-594,
 * src/stim/py/compiled_measurement_sampler.pybind.cc:27-80) and sample_batch_measurements
 * (src/stim/simulators/frame_simulator_util.inl:290-372). */
int mfs_compile_measurements(const char *stim_text, size_t len, uint64_t seed, int has_seed, int skip_reference_sample,
                             const uint8_t *reference_sample_bits, size_t reference_len, mfs_sampler **out);

/* Measurement-to-detection-event converter. Compiles the circuit's detectors and observables, its
 * reference sample (all zeros with skip_reference_sample = 1) and, for circuits with sweep-controlled
 * gates, each sweep bit's effect on them.
 * Reference: Circuit.compile_m2d_converter (src/stim/circuit/circuit.pybind.cc:781-785) ->
 * CompiledMeasurementsToDetectionEventsConverter (src/stim/simulators/measurements_to_detection_events.pybind.cc:159-166)
 * and the CLI's `stim m2d` (src/stim/cmd/command_m2d.cc). */
typedef struct mfs_m2d mfs_m2d;
int mfs_m2d_compile(const char *stim_text, size_t len, int skip_reference_sample, mfs_m2d **out);
void mfs_m2d_destroy(mfs_m2d *c);
/* out[0..3] = num_measurements, num_detectors, num_observables, num_sweep_bits (largest sweep[k]
 * index plus one). */
int mfs_m2d_stats(const mfs_m2d *c, uint64_t out[4]);
/* Converts `shots` bit-packed little-endian measurement rows (meas_stride bytes apart) into detection
 * event rows (with observables appended when MFS_APPEND_OBSERVABLES) and, with
 * MFS_SEPARATE_OBSERVABLES, observable rows; MFS_BIT_PACKED selects packed outputs (else one byte per
 * entry); MFS_DEVICE_OUTPUT means every pointer is device memory. Each detector (observable) is the
 * XOR of its measurements, inverted when the XOR of their reference bits is 1. Packed device
 * measurement rows with 16-byte-aligned address and stride >= 16 * ceil(M / 128) are read in place,
 * 16 bytes at a time, so that buffer must span shots * meas_stride bytes. Output strides must be
 * below 2^30 bytes.
 * Reference: CompiledMeasurementsToDetectionEventsConverter.convert
 * (measurements_to_detection_events.pybind.cc:78-157) and measurements_to_detection_events_helper
 * (src/stim/simulators/measurements_to_detection_events.inl:29-131). */
int mfs_m2d_convert(mfs_m2d *c, uint64_t shots, const uint8_t *meas, size_t meas_stride, uint32_t flags, uint8_t *dets,
                    size_t dets_stride, uint8_t *obs, size_t obs_stride);
/* mfs_m2d_convert with sweep data: sweep holds `shots` bit-packed little-endian rows of
 * ceil(num_sweep_bits / 8) bytes, sweep_stride bytes apart (device memory with MFS_DEVICE_OUTPUT), or
 * is NULL for all-zero sweep bits. Each sweep bit flips the detectors and observables that the
 * noiseless circuit's gates controlled by it reach (tabulated when the converter is compiled).
 * Reference: convert(..., sweep_bits=...) and convert_file(..., sweep_bits_filepath, sweep_bits_format)
 * (measurements_to_detection_events.pybind.cc:44-157), `stim m2d --sweep` (command_m2d.cc:21-75), and the
 * sweep-driven frame simulator of measurements_to_detection_events.inl:57-75, 108-114. */
int mfs_m2d_convert_sweep(mfs_m2d *c, uint64_t shots, const uint8_t *meas, size_t meas_stride, const uint8_t *sweep,
                          size_t sweep_stride, uint32_t flags, uint8_t *dets, size_t dets_stride, uint8_t *obs,
                          size_t obs_stride);
/* Benchmark helper (no reference counterpart): average kernel time in ms of converting `shots`
 * device-resident rows (observables appended), measured with HIP events on the converter's stream. */
int mfs_m2d_bench(mfs_m2d *c, uint64_t shots, int warmup, int iters, double out_ms[1]);
/* Kernel selection (no reference counterpart; results are identical either way). Circuits without
 * sweep bits whose detectors only reach back a bounded number of results (memory experiments: one
 * round) can be converted by a windowed kernel that keeps a ring of at most 8,192 measurement words
 * per 32-shot tile in LDS instead of the whole record. kernel: 0 = windowed where it applies and needs
 * less than half the LDS (default), 1 = whole-record kernel, 2 = windowed (conversions it does not
 * apply to then fail with MFS_ERR_INVALID_ARGUMENT). */
int mfs_m2d_set_kernel(mfs_m2d *c, int kernel);
/* out[0..3] = ring words of the windowed kernel per output layout (0: not applicable): observables
 * appended, observables separate, detectors only, observables appended and separate; out[4] = LDS
 * bytes per tile of the whole-record kernel, out[5] = of the windowed kernel (appended layout; 0 when
 * not applicable). */
int mfs_m2d_kernel_info(const mfs_m2d *c, uint64_t out[6]);

/* Reference sample of a circuit (host only, no GPU): the noiseless run with every random measurement
 * result taken as 0, written bit-packed little-endian into out[0 .. ceil(num_measurements / 8)).
 * Reference: Circuit.reference_sample (src/stim/circuit/circuit.pybind.cc:664-671) and
 * TableauSimulator::reference_sample_circuit (src/stim/simulators/tableau_simulator.inl:1435-1438). */
int mfs_reference_sample(const char *stim_text, size_t len, uint8_t *out, size_t out_len);

/* Host only, no GPU (no reference counterpart; the measurement sampler's expansion path, which stands
 * in for the frame simulation of sample_batch_measurements, frame_simulator_util.inl:290-372):
 * out[0..7] = whether the expansion applies (1) or not (0), measurements, pivots (independent
 * detector/observable rows), unconstrained results, random results and random collapses of the
 * reference run, substitution phases (barriers of the expansion kernel's schedule), substitution
 * depth (longest dependency chain). */
int mfs_expand_plan_stats(const char *stim_text, size_t len, uint64_t out[8]);

/* Host only, a test hook (no reference counterpart): the expansion of `shots` shots computed on the
 * host as the kernel computes it, one byte per bit: y rows hold the detector then observable values,
 * free rows one value per unconstrained result (in record order), out rows the measurement results
 * before reference inversion. Fails when the expansion does not apply. */
int mfs_expand_shots_host(const char *stim_text, size_t len, uint64_t shots, const uint8_t *y, size_t y_stride,
                          const uint8_t *free_bits, size_t free_stride, uint8_t *out, size_t out_stride);

/* Samples detection events. Row strides are in bytes. dets must hold shots rows of
 * (packed ? ceil(n/8) : n) bytes where n = D + O*(prepend+append); obs (when
 * MFS_SEPARATE_OBSERVABLES) holds shots rows of (packed ? ceil(O/8) : O) bytes.
 * Reference: CompiledDetectorSampler::sample_to_numpy (compiled_detector_sampler.pybind.cc:34-86). */
int mfs_sample(mfs_sampler *s, uint64_t shots, uint32_t flags, uint8_t *dets, size_t dets_stride, uint8_t *obs,
               size_t obs_stride);

/* Allocation-free form of mfs_sample for device-resident rows (MFS_DEVICE_OUTPUT required). No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). It calls no hipMalloc, hipFree
 * or allocating library routine: device staging (packed rows before they are expanded to one byte per
 * entry, unpacked rows before packing across a non-byte-aligned observable boundary) comes from a fixed
 * 64 MiB region that each GPU holds for the life of the process, and calls are cut into batches of
 * whole RNG tiles that fit it, so the rows equal mfs_sample's bit for bit. A call whose rows do not fit
 * the region one tile at a time fails with MFS_ERR_INVALID_ARGUMENT, and so does a call that would
 * first have to build the sampler's symptom table (prepare the sampler once with mfs_engine_info or an
 * allocating call). Calls that overlap on one GPU share the region and race. */
int mfs_sample_noalloc(mfs_sampler *s, uint64_t shots, uint32_t flags, uint8_t *dets, size_t dets_stride, uint8_t *obs,
                       size_t obs_stride);

/* Samples and writes to a file in format "01", "b8", "r8", "ptb64", "hits" or "dets".
 * obs_path may be NULL. prepend + append + (obs_path != NULL) > 1 -> MFS_ERR_OUT_OF_RANGE.
 * Reference: sample_batch_detection_events_writing_results_to_disk (frame_simulator_util.inl:114-187, 206-288). */
int mfs_sample_write(mfs_sampler *s, uint64_t shots, const char *path, const char *format, int prepend_observables,
                     int append_observables, const char *obs_path, const char *obs_format);

/* Fixed-fault mode (no reference counterpart; the reference's equivalent is FlipSimulator with
 * disable_stabilizer_randomization=True plus broadcast_pauli_errors, frame_simulator.pybind.cc:310-329,
 * 1365-1421): no noise is drawn; shot s gets exactly the faults faults[3*i .. 3*i+2] for
 * i in [shot_offsets[s], shot_offsets[s+1]), each (instruction index in the REPEAT-unrolled circuit,
 * site index inside that instruction, outcome code). Codes: single-qubit channels 1=X 2=Z 3=Y;
 * two-qubit channels bits (x1, z1, x2, z2); noisy results 1=flip; E/ELSE 1=fires; heralded
 * channels bit2 = herald, bits 0/1 = X/Z. */
int mfs_sample_with_faults(mfs_sampler *s, uint64_t shots, const uint64_t *shot_offsets, const uint32_t *faults,
                           uint32_t flags, uint8_t *dets, size_t dets_stride, uint8_t *obs, size_t obs_stride);

/* Benchmark helper: runs warmup + iters samples of `shots` into device-resident packed outputs
 * (separate observables) and reports average times in milliseconds measured with HIP events on the
 * sampler's stream: out_ms[0] = the whole sampling step, out_ms[1] = its output-writing kernel
 * (the same single kernel on every current path). */
int mfs_bench_device(mfs_sampler *s, uint64_t shots, int warmup, int iters, double out_ms[2]);

/* Sampling engine: 0 = automatic, 1 = frame engine (the reference's per-instruction frame
 * simulation, FrameSimulator::do_circuit, frame_simulator.inl:165-170), 2 = symptom engine
 * (frame-kernel symptoms of every basis fault + sparse noise sampling; same distribution,
 * deterministic-detector circuits only). Automatic picks the symptom engine when it applies.
 * For a measurement sampler, 2 = expansion: detection events and observable flips drawn on the
 * symptom engine, expanded into measurement records by forward substitution with the unconstrained
 * (random) results drawn uniformly; same distribution, for circuits whose detectors and observables
 * fix every parity the noiseless circuit fixes and whose random collapses are all recorded.
 * No reference counterpart (the reference has one engine). */
int mfs_set_engine(mfs_sampler *s, int engine);

/* out[0..9] = engine in use for sampling (1 frame, 2 symptom, 3 expansion), site classes, symptom
 * columns, symptom entries, symptom-table build time in microseconds, noise sites, for the symptom
 * engine the shots per workgroup tile and the sampling kernel's LDS bytes per tile (bit-packed output
 * with separate observables), and the specialised scan kernel's distinct outcome blocks and outcome
 * records (0 when it does not apply). A measurement sampler on the expansion path (3) reports its
 * detector sampler's symptom figures in out[1..5] and, in out[6..8], its free (random) results,
 * substitution phases and substitution items. Builds the symptom table or expansion if needed. */
int mfs_engine_info(mfs_sampler *s, uint64_t out[10]);

/* Symptom engine: target size (KB) of one workgroup's LDS image of output rows (sets the shots per
 * workgroup). Default 32. A negative value selects the per-lane-walk queue kernel instead of the
 * prefix-scan kernel, with image size -kb. */
int mfs_set_symptom_tile_kb(mfs_sampler *s, int kb);

/* ---- Detector error model sampler -------------------------------------------------------------
 * Replaces stim.DetectorErrorModel.compile_sampler(seed=...) and CompiledDemSampler.sample / sample_write
 * (src/stim/dem/detector_error_model.pybind.cc, src/stim/simulators/dem_sampler.pybind.cc:40-79, 119-309;
 * the sampler itself src/stim/simulators/dem_sampler.inl:50-77). The model crosses as .dem text. The
 * returned handle works with mfs_sample, mfs_sample_write, mfs_bench_device, mfs_circuit_stats
 * (qubits, measurements and lookback are 0) and mfs_destroy. mfs_sample_with_faults replays recorded
 * errors (recorded_errors_to_replay): a fault (e, 0, 1) fires mechanism e, counted over error
 * instructions in REPEAT-unrolled order. Every mechanism fires independently with its own probability. */
int mfs_dem_compile(const char *dem_text, size_t len, uint64_t seed, int has_seed, mfs_sampler **out);
/* Samples like mfs_sample (observables separate) and also returns which mechanisms fired: errs holds
 * `shots` rows of num_errors entries (bit-packed little-endian with MFS_BIT_PACKED, else one byte per
 * mechanism), errs_stride bytes apart, mechanism e counted over error instructions in REPEAT-unrolled
 * order. Reference: CompiledDemSampler.sample(shots, bit_packed, return_errors=True) and the err_out
 * file of sample_write (dem_sampler.pybind.cc:45-79, dem_sampler.inl:75-99, where err_buffer holds each
 * mechanism's fired bits). */
int mfs_dem_sample_errors(mfs_sampler *s, uint64_t shots, uint32_t flags, uint8_t *dets, size_t dets_stride,
                          uint8_t *obs, size_t obs_stride, uint8_t *errs, size_t errs_stride);
/* Parses on the host only: out[0..2] = detectors, observables, error mechanisms. */
int mfs_dem_parse_stats(const char *dem_text, size_t len, uint64_t out[3]);

/* ---- Batches: many sampling jobs in one persistent kernel ----------------------------------------
 * No reference counterpart (the reference samples one circuit per call). A batch holds jobs, each a
 * sampler, a shot count and device-resident bit-packed output rows (dets: shots rows of >= ceil(D/8)
 * bytes at dets_stride; obs: shots rows of >= ceil(O/8) bytes at obs_stride, or NULL to drop the
 * observables). Jobs may come from different circuits. A run samples every job: symptom-engine jobs in
 * one persistent kernel that hands out shot tiles from a work counter, heaviest tiles first; other
 * jobs as their own launches. Each job draws exactly what mfs_sample(s, shots, MFS_BIT_PACKED |
 * MFS_SEPARATE_OBSERVABLES | MFS_DEVICE_OUTPUT, ...) would at the same point of its sampler's stream
 * with the same tile size, and advances the sampler's stream the same way. A batch and its samplers
 * are single-threaded; samplers must outlive the batch. */
typedef struct mfs_batch mfs_batch;
int mfs_batch_create(mfs_batch **out);
void mfs_batch_destroy(mfs_batch *b);
int mfs_batch_add(mfs_batch *b, mfs_sampler *s, uint64_t shots, uint8_t *dets, size_t dets_stride, uint8_t *obs,
                  size_t obs_stride);
/* LDS output image per workgroup in KB for every job (0 = each sampler's own setting). */
int mfs_batch_set_tile_kb(mfs_batch *b, int kb);
/* Runs every job once and waits for the results. use_graph != 0 captures the persistent kernel and its
 * bookkeeping kernels into a HIP graph on the first run and replays it afterwards. */
int mfs_batch_run(mfs_batch *b, int use_graph);
/* Runs warmup + iters times; out_ms[0] = average device time per run (HIP events on the batch's
 * stream), out_ms[1] = average wall time per run. */
int mfs_batch_time(mfs_batch *b, int use_graph, int warmup, int iters, double out_ms[2]);
/* Schedule of the batch kernel: 1 (default) = persistent workgroups (one per resident slot) claiming
 * chunks of tiles from a work counter; 0 = one workgroup per tile. Results do not depend on it. */
int mfs_batch_set_schedule(mfs_batch *b, int persistent);
/* Autotuner: times iters runs for each candidate tile size (KB) under both schedules and keeps the
 * fastest pair, also setting the tile size on every job's sampler so plain calls stay comparable. */
int mfs_batch_autotune(mfs_batch *b, const int *candidates, int n, int iters, int *best_kb, double *best_ms);
/* out[0..7] = jobs, kernel jobs, separately launched jobs, tiles, workgroups, LDS bytes, persistent
 * schedule (1/0), tile size in KB (0 = each sampler's own). */
int mfs_batch_info(mfs_batch *b, uint64_t out[8]);

/* ---- Serving: resident circuits, memory-budget admission, continuous batching -------------------
 * No reference counterpart. A server takes requests (sampler, shots, request seed) for circuits with
 * deterministic detectors (symptom engine). A request's bit-packed rows are allocated in device memory
 * when it is admitted, which happens in arrival order as soon as they fit budget_bytes; they stay
 * until mfs_server_release. Each mfs_server_step admits what fits, takes a fair share of every running
 * request's shot tiles (about step_cost_bytes of estimated work in total: output bytes plus 16 bytes
 * per sampled fault), runs them in one persistent-kernel launch and waits. A request's result depends
 * only on its sampler, seed and shot count, never on the other requests or on how its tiles were
 * split across steps. Single-threaded; samplers must outlive their requests. */
typedef struct mfs_server mfs_server;
int mfs_server_create(uint64_t budget_bytes, double step_cost_bytes, mfs_server **out);
void mfs_server_destroy(mfs_server *v);
int mfs_server_submit(mfs_server *v, mfs_sampler *s, uint64_t shots, uint64_t seed, uint64_t *out_id);
/* *did_work = 0 when nothing was queued or running. */
int mfs_server_step(mfs_server *v, int *did_work);
/* Moves up to cap finished request ids into ids; *n = how many. */
int mfs_server_poll(mfs_server *v, uint64_t *ids, uint32_t cap, uint32_t *n);
/* Device pointers and strides of a finished request's rows; times = submitted, admitted, finished
 * (seconds, steady clock). */
int mfs_server_result(mfs_server *v, uint64_t id, uint8_t **dets, size_t *dets_stride, uint8_t **obs,
                      size_t *obs_stride, uint64_t *shots, double times[3]);
/* Copies a finished request's rows to host buffers. */
int mfs_server_copy_result(mfs_server *v, uint64_t id, uint8_t *dets, size_t dets_stride, uint8_t *obs,
                           size_t obs_stride);
/* Frees a finished (or still queued) request and its rows. */
int mfs_server_release(mfs_server *v, uint64_t id);
/* out[0..7] = steps, queued, running, bytes in use, peak bytes in use, budget, device time of the
 * serving launches in microseconds (HIP events), row bytes those launches wrote. */
int mfs_server_stats(mfs_server *v, uint64_t out[8]);
/* Replays a request trace in real time: request i (samplers[i], shots[i], seeds[i]) arrives
 * arrival_s[i] seconds after the call starts (sorted); each result is released as soon as it is
 * complete. latency_ms[i] = finish - arrival; *wall_s = duration of the whole replay. */
int mfs_server_replay(mfs_server *v, uint64_t n, mfs_sampler *const *samplers, const uint64_t *shots,
                      const uint64_t *seeds, const double *arrival_s, double *latency_ms, double *wall_s);

/* Configuration knobs: words per tile (1, 2 or 4; 0 = automatic). Returns the value in use. */
int mfs_set_tile_words(mfs_sampler *s, int words);
int mfs_get_info(const mfs_sampler *s, uint64_t out[8]); /* tile_words, lds_bytes, ops, syncs, noise_ops, randomize, ring, dense_qubits */

/* Symptom-engine scan kernel: 0 (default) = the specialised kernel with one precomputed outcome record
 * per (noise site, outcome) where the model allows it (plain groups with uniform outcomes, as in the
 * QEC memory circuits), else the generic kernel; 1 = always the generic kernel, which applies a hit's
 * basis-column records one by one. Both draw the same hits, so rows are bit-identical. No reference
 * counterpart. */
int mfs_set_scan_variant(mfs_sampler *s, int variant);

/* Position of a sampler's random stream: the global shot-tile index its next call starts at. Every
 * call advances it by the tiles it used, so consecutive calls draw fresh randomness; setting it back
 * reproduces a call. Tile indices are 64-bit and the stream does not repeat. No reference counterpart
 * (the reference's samplers hold one std::mt19937_64, src/stim/util_bot/probability_util.h:40-76). */
int mfs_set_stream_position(mfs_sampler *s, uint64_t tiles);
int mfs_get_stream_position(const mfs_sampler *s, uint64_t *tiles);

/* Checks of the random streams (no reference counterpart; the reference's generator is libstdc++'s
 * std::mt19937_64). mfs_rng_counter reports the Philox counter out[0..3] and key out[4..5] that the
 * symptom kernels (kernel 0) or the frame kernel (kernel 1) use for draw `draw` of lane `lane` in
 * shot tile `tile`, stream tag `tag`, under kernel key `key` (host only). mfs_philox runs the
 * kernels' Philox4x32-10 on the GPU over n counters (4 words each) under one key. */
int mfs_rng_counter(int kernel, uint64_t key, uint64_t tile, uint32_t lane, uint32_t draw, uint32_t tag,
                    uint32_t out[6]);
int mfs_philox(const uint32_t *counters, uint64_t n, uint64_t key, uint32_t *out);

/* Hash of the sources the library was built from (16 hex digits; stim_amd/_buildid.py). */
const char *mfs_build_id(void);

const char *mfs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
