// Measurement-to-detection-event conversion: host structures shared with the kernel.
//
// Reference: measurements_to_detection_events_helper (src/stim/simulators/measurements_to_detection_events.inl:29-131).
// Detector d of a shot is the XOR of the shot's measurement results that d names, inverted when the
// XOR of those measurements' reference-sample bits is 1, XOR the shot's sweep bits that reach d;
// observable k likewise over the measurement records OBSERVABLE_INCLUDE(k) names. A sweep bit reaches
// an output through the gates it controls (CX/CY/CZ/XCZ/YCZ with a sweep[k] control flip a qubit's
// frame; the reference tracks them with a noiseless FrameSimulator driven by the sweep table,
// .inl:57-75): the noiseless circuit is linear in those flips, so each sweep bit's effect on the
// outputs is a fixed set of columns, computed once when the converter is built.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sa {

struct M2dParams {
    const uint8_t *in;          // bit-packed measurement rows (shot-major), M bits each
    uint64_t in_stride;
    uint32_t in_row_bytes;      // ceil(M / 8)
    uint32_t in_vec;            // rows are 16-byte aligned with 16-byte-multiple strides
    uint32_t blocks;            // ceil(M / 32)
    uint32_t num_det;           // D: columns 0 .. D-1 are detectors, D .. D+O-1 observables
    uint8_t *out;               // bit-packed output rows: n_out columns starting at column 0
    uint64_t out_stride;
    uint32_t n_out;
    uint32_t n_obs;             // separate observable rows (columns D .. D+n_obs-1), or 0
    uint8_t *obs;
    uint64_t obs_stride;
    uint64_t shots;
    const uint2 *cols;          // column records over the D + O columns (below)
    const uint32_t *ovf;        // overflow term lists: count, then absolute measurement indices
    const uint8_t *sweep;       // bit-packed sweep-bit rows (shot-major), or null: every sweep bit 0
    uint64_t sweep_stride;
    uint32_t sweep_row_bytes;   // ceil(K / 8) for the circuit's K sweep bits
    uint32_t sweep_blocks;      // ceil(K / 32): LDS word m2d_sweep_word(blocks) + k holds sweep bit k
};

// Expansion of detection events into measurement records (the measurement sampler's fast path; plan
// and table layouts in expand.h). One workgroup expands a tile of 32 shots: its measurement words start
// at zero, free results get random words, each y column's word is XORed into its slots, the
// substitution phases run (each thread its own items, a barrier between phases), and the words leave
// as the columns of `out` (column m = word m, inverted where the reference sample is 1: bit m % 32 of
// inv[m / 32]).
// Substitution records of the expansion: 8 words (two uint4) each, up to 7 terms inline (layout in
// expand.h).
constexpr uint32_t EXPAND_RECORD_WORDS = 8;
constexpr uint32_t EXPAND_INLINE_TERMS = 7;

struct ExpandDev {
    const uint8_t *in;          // y rows: detection events then observable flips, bit packed; 16-byte
    uint64_t in_stride;         // aligned rows with a stride that is a multiple of 16
    uint32_t in_groups;         // 16-byte groups of a y row that hold its Y bits
    uint32_t Y;
    const uint32_t *prim, *extra_off, *extra;
    const uint32_t *free_slots;
    uint32_t nfree;
    uint32_t nphases;
    const uint32_t *sched_off;  // nphases * 1024 + 1 record offsets (phase-major, then thread)
    const uint4 *items;         // two uint4 per record (8 words, layout in expand.h)
    const uint32_t *terms;
    uint64_t seed;              // Philox key of the free results
    uint64_t tile0;             // 32-shot tile of the launch's first shot in the sampler's stream
    const uint32_t *inv;        // ceil(M / 32) inversion words, zero past the last measurement
    M2dParams out;              // output rows (out, out_stride, n_out = M, shots, blocks)
};

// Column record: bits 28..29 of x = term count class c, bit 31 = the column's reference parity
// (invert). c = 0..2: the column has c terms; x bits 0..27 = first measurement index, y = second
// (unused fields name the always-zero word, m2d_zero_word, so both words are read and XORed). c = 3: x bits 0..27 = term count n, y = offset of the term list in the
// overflow table (a multiple of 4); the list holds the n indices padded to a multiple of 4 with
// the index of a word that is always zero (m2d_zero_word), so it is read 16 bytes at a time.
constexpr uint32_t M2D_INDEX_MASK = 0x0FFFFFFFu;
constexpr uint32_t M2D_INVERT = 1u << 31;

// Windowed conversion (circuits whose columns reach back a bounded number of results, the memory
// experiments among them): a 256-lane workgroup streams its tile's measurements through a ring of
// ring_words words (a power of two) a chunk of M2D_WIN_CHUNK measurements at a time, and emits each
// block of 32 output columns once its last term has passed. Output blocks of part 0 (columns
// [0, n_out)) are emitted in order: blocks [chunk_end[ch - 1], chunk_end[ch]) after chunk ch is loaded;
// part 1 (the n_obs separate observable columns) after the last chunk. The host checks that every
// block's terms are still in the ring when it is emitted. Without sweep bits only.
constexpr uint32_t M2D_WIN_CHUNK = 1024;
constexpr uint32_t M2D_WIN_MAX_CHUNKS = 1024;   // chunk_end is copied to LDS
struct M2dWindow {
    uint32_t ring_words;
    uint32_t nchunks;           // ceil(M / M2D_WIN_CHUNK), at most M2D_WIN_MAX_CHUNKS
    const uint32_t *chunk_end;  // per chunk: part-0 blocks emitted once it is loaded (non-decreasing)
    uint32_t zero_rec;          // index in cols of a zero record (no terms, no inversion)
};

// LDS words of a tile: measurement blocks rounded up to whole 16-byte load groups (4 blocks of 32
// words), then the sweep-bit blocks, then 4 words of zeros. Column terms are word indices: measurement
// m is word m, sweep bit k is word m2d_sweep_word(blocks) + k.
__host__ __device__ inline uint32_t m2d_word_blocks(uint32_t blocks) {
    return (blocks + 3) / 4 * 4;
}
__host__ __device__ inline uint32_t m2d_sweep_word(uint32_t blocks) {
    return m2d_word_blocks(blocks) * 32;
}
__host__ __device__ inline uint32_t m2d_zero_word(uint32_t blocks, uint32_t sweep_blocks) {
    return m2d_sweep_word(blocks) + sweep_blocks * 32;
}

// Dynamic LDS of a conversion or expansion launch: the tile's words plus the staged output chunk.
size_t m2d_lds_bytes(uint32_t blocks, uint32_t sweep_blocks);
// Dynamic LDS of a windowed launch: the ring, the staged output blocks and the chunk_end table.
size_t m2d_win_lds_bytes(uint32_t ring_words, uint32_t nchunks);

}  // namespace sa
