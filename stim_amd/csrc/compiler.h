// Lowers a flattened circuit into the device program executed by the frame kernel.
#pragma once

#include <cstdint>
#include <vector>

#include "circuit.h"
#include "program.h"

namespace sa {

// Where the sites of one flat instruction landed in the device program (one entry per emitted
// noise op; an instruction split into segments has several entries).
struct SiteRange {
    uint32_t op;        // device op index
    uint32_t first;     // first site (group index within the instruction) covered by the op
    uint32_t count;     // sites covered
};

struct Program {
    std::vector<DevOp> ops;
    std::vector<DevEpoch> epochs;
    std::vector<uint32_t> tgt;    // frame-op words (before packing into blobs)
    std::vector<uint32_t> blob;   // epoch blobs (program.h layout) as the device walks them
    std::vector<uint32_t> ntgt;   // noise-op words
    std::vector<uint32_t> otab;
    uint32_t num_qubits = 0;      // dense qubit rows
    uint32_t ring = 1;            // record ring rows
    uint32_t blob0_len = 0;       // words of the first epoch blob
    uint32_t num_detectors = 0;
    uint32_t num_observables = 0;
    uint64_t num_measurements = 0;
    uint32_t num_noise_ops = 0;
    uint32_t num_syncs = 0;
    bool has_randomizing_ops = false;
    // flat instruction index -> site ranges of its noise ops
    std::vector<std::vector<SiteRange>> sites;
    Stats stats;
};

// measurement_outputs: the outputs are the circuit's measurement results in record order (one
// output row per result) instead of its detectors and observables.
Program compile_program(const std::vector<const Instr *> &flat, bool measurement_outputs = false);

// Resolves a fault (flat instruction index, group index, code) to (op, site); throws
// std::out_of_range when the instruction has no such noise site.
void resolve_fault(const Program &prog, uint32_t k, uint32_t j, uint32_t *op, uint32_t *site);

// LDS bytes the frame kernel needs for a tile of 32*T shots.
size_t lds_bytes(const Program &prog, int T);

}  // namespace sa
