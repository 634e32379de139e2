// Device program for the frame-simulation kernel: a flat list of ops that every workgroup walks
// in order over its own shot tile. Shared by the host compiler (program.cc) and the HIP kernels.
//
// Each op is one of the reference's FrameSimulator do_* handlers (src/stim/simulators/
// frame_simulator.inl:139-1113) restated for a tile of 32*T shots held in LDS, with its targets
// split so that no two items of one op touch the same qubit / record slot (the reference applies
// an instruction's targets one after another; splitting keeps that order where it matters).
#pragma once

#include <cstdint>

namespace sa {

enum OpCode : uint32_t {
    OP_NOP = 0,
    // Single-qubit Clifford frame rules (items = qubits).
    OP_H,        // do_H_XZ: swap(x, z)
    OP_HXY,      // do_H_XY: z ^= x
    OP_HYZ,      // do_H_YZ: x ^= z
    OP_CXYZ,     // do_C_XYZ: x ^= z; z ^= x
    OP_CZYX,     // do_C_ZYX: z ^= x; x ^= z
    // Two-qubit Clifford frame rules (items = pairs).
    OP_CX, OP_CY, OP_CZ, OP_SWAP, OP_ISWAP, OP_CXSWAP, OP_CZSWAP, OP_SWAPCX,
    OP_SQRTXX, OP_SQRTYY, OP_SQRTZZ, OP_XCX, OP_XCY, OP_YCX, OP_YCY,
    // Classical control (items = (record slot, qubit, mask)): x ^= rec if mask&1, z ^= rec if mask&2.
    OP_FEEDBACK,
    // Resets (items = qubits). Randomization flag fills the commuting row with 50% bits.
    OP_R, OP_RX, OP_RY,
    // Single-qubit measurements (items = qubits; record slot = base + i).
    OP_M, OP_MX, OP_MY, OP_MR, OP_MRX, OP_MRY,
    // Pauli-product measurement (items = products, CSR terms; record slot = base + i).
    OP_MPP,
    // Pauli-product rotation (items = products, CSR terms with optional record terms).
    OP_SPP,
    // Zero record slots [base, base + n) (MPAD results, herald records before sampling).
    OP_RECZERO,
    // Detectors (items = detectors, CSR record-slot terms; staging row = (base + i) % STAGE_ROWS).
    OP_DET,
    // Observables (items = (index, CSR terms: record slots or Pauli frame terms)).
    OP_OBS,
    // Write staging rows [0, n) as detectors base .. base + n to the output.
    OP_FLUSH,
    // Write observable rows to the output.
    OP_FLUSH_OBS,
    // Noise channels. Sites: qubits (NOISE1), pairs (NOISE2), record slots (MFLIP),
    // one Pauli product (CORR), qubits + record slots (HERALD).
    OP_NOISE1, OP_NOISE2, OP_MFLIP, OP_CHAINCLR, OP_CORR, OP_HERALD,
    OP_COUNT
};

// Flag bits in DevOp::code above the low 8 opcode bits.
constexpr uint32_t OPF_SYNC = 1u << 8;        // barrier after this op
constexpr uint32_t OPF_RANDOMIZE = 1u << 9;   // frame randomization (R/M/MR/MPP) when enabled
constexpr uint32_t OPF_CHAIN_START = 1u << 10; // OP_CORR that starts a chain (E)
constexpr uint32_t OPF_DENSE = 1u << 11;      // noise drawn densely (high probability)

// Term encodings inside CSR lists.
constexpr uint32_t TERM_PAULI = 1u << 31;   // frame term: qubit | X<<24 | Z<<25
constexpr uint32_t TERM_X = 1u << 24;
constexpr uint32_t TERM_Z = 1u << 25;
constexpr uint32_t TERM_QMASK = (1u << 24) - 1;

struct DevOp {
    uint32_t code;   // opcode | flags
    uint32_t n;      // items (sites for noise ops)
    uint32_t toff;   // frame ops: word offset in the epoch blob; noise ops: offset into ntgt
    uint32_t base;   // record slot base / detector base / term count (CORR)
    float p;         // total firing probability (noise)
    float inv_log;   // 1 / log2(1 - p) for geometric gaps (noise)
    uint32_t woff;   // offset into the outcome table (pairs of (threshold, code))
    uint32_t nout;   // number of outcomes (noise) | chunk log2 << 16
};
static_assert(sizeof(DevOp) == 32, "DevOp layout");

// A run of ops with no barrier inside (host bookkeeping).
struct DevEpoch {
    uint32_t op_begin, op_end;
    uint32_t tbegin, tlen;
};

// The device walks the program as a chain of epoch blobs. Blob layout (32-bit words):
//   [0] ops in this epoch  [1] word offset of the next blob  [2] its length (0: last)
//   [3] global index of this epoch's first op  [4..7] unused
//   [8 .. 8 + 8*ops)        DevOp headers
//   [8 + 8*ops .. len)      the frame ops' words (targets, CSR lists); DevOp::toff is relative
//                           to the blob start.
// Each blob is copied into one of two LDS staging buffers while the previous epoch runs, so the
// op loop reads headers and targets from LDS only.
// LDS one workgroup may use on gfx950.
constexpr uint32_t LDS_PER_WORKGROUP = 160 * 1024;

constexpr uint32_t BLOB_HEADER = 8;
constexpr uint32_t STAGE_WORDS = 1024;  // staged words per epoch blob (4 KB; double-buffered)
constexpr uint32_t OP_WORDS_MAX = STAGE_WORDS - 2 * BLOB_HEADER;  // words of one frame op

constexpr uint32_t STAGE_ROWS = 256;   // detectors staged per flush

// One injected fault in fixed-fault mode, pre-sorted by (tile, op).
struct DevFault {
    uint32_t op;     // op index of the noise op
    uint32_t site;   // site within the op
    uint32_t shot;   // shot within the tile
    uint32_t code;   // outcome code (see FrameOracle / frame_oracle.py fault addressing)
};

// Output description for one destination (detector rows or observable rows).
struct DevOut {
    uint8_t *ptr;       // device pointer to row 0 of this launch
    uint64_t stride;    // bytes per shot row
    uint32_t col0;      // bit (packed) or byte (unpacked) column of entry 0
    uint32_t packed;    // 1: bit-packed little-endian; 0: one byte per entry
    uint32_t base_char; // unpacked: value added to each bit (0 for bool, '0' for text)
    uint32_t enabled;
};

struct LaunchParams {
    const DevFault *faults;    // fixed-fault mode: all faults of this launch
    const uint32_t *fault_off; // per tile start offsets (num_tiles + 1)
    DevOut dets;
    DevOut obs;
    uint64_t seed;
    uint64_t shots;            // valid shots in this launch
    uint64_t tile0;            // global tile index of block 0 (RNG stream)
    uint32_t blob0_len;        // length of the first epoch blob (it starts at word 0)
    uint32_t num_qubits;       // dense qubit rows
    uint32_t ring;             // record ring rows
    uint32_t num_obs;
    uint32_t mode;             // 0 = no noise, 1 = sampled, 2 = fixed faults
    uint32_t randomize;        // 1 = frame randomization
    const uint32_t *det_invert;  // optional bit per output row: the row is inverted (reference sample bits)
};

// Device-resident program buffers.
struct DevProgram {
    const uint32_t *blob;      // epoch blobs
    const uint32_t *ntgt;      // noise-op words (read on hits)
    const uint32_t *otab;      // outcome table (threshold, code) pairs
};

constexpr uint32_t MODE_NOISELESS = 0;
constexpr uint32_t MODE_SAMPLED = 1;
constexpr uint32_t MODE_FIXED = 2;

}  // namespace sa
