// Circuit text (.stim) parser and REPEAT flattening for the stim_amd sampler host.
//
// Boundary input: the circuit crosses the C-ABI as .stim text (doc/file_format_stim_circuit.md).
// Gate names, aliases and target shapes follow src/stim/gates/gates.h:93-196 and
// src/stim/gates/gate_data_*.cc; target kinds follow src/stim/circuit/gate_target.h:27-33;
// execution order of REPEAT blocks follows Circuit::for_each_operation (src/stim/circuit/circuit.h:181-193).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sa {

enum class Gate : uint8_t {
    DETECTOR, OBSERVABLE_INCLUDE, TICK, QUBIT_COORDS, SHIFT_COORDS, MPAD,
    MX, MY, M, MRX, MRY, MR, RX, RY, R,
    XCX, XCY, XCZ, YCX, YCY, YCZ, CX, CY, CZ,
    H, H_XY, H_YZ, H_NXY, H_NXZ, H_NYZ,
    DEPOLARIZE1, DEPOLARIZE2, X_ERROR, Y_ERROR, Z_ERROR, I_ERROR, II_ERROR,
    PAULI_CHANNEL_1, PAULI_CHANNEL_2, E, ELSE_CORRELATED_ERROR,
    HERALDED_ERASE, HERALDED_PAULI_CHANNEL_1,
    I, X, Y, Z,
    C_XYZ, C_NXYZ, C_XNYZ, C_XYNZ, C_ZYX, C_ZYNX, C_ZNYX, C_NZYX,
    SQRT_X, SQRT_X_DAG, SQRT_Y, SQRT_Y_DAG, S, S_DAG,
    II, SQRT_XX, SQRT_XX_DAG, SQRT_YY, SQRT_YY_DAG, SQRT_ZZ, SQRT_ZZ_DAG,
    SWAP, ISWAP, ISWAP_DAG, CXSWAP, SWAPCX, CZSWAP,
    MXX, MYY, MZZ, MPP, SPP, SPP_DAG,
    NUM_GATES
};

// Shape of the targets a gate takes.
enum class Shape : uint8_t { NONE, QUBITS, PAIRS, PAULI_PRODUCTS, RECORDS, ANNOTATION_ANY };

struct GateInfo {
    const char *name;
    Gate gate;
    Shape shape;
    int8_t min_args;  // -1: any
    int8_t max_args;  // -1: any
    bool can_target_bits;   // CX/CY/CZ/XCZ/YCZ accept rec/sweep controls
    bool produces_results;
};

const GateInfo &gate_info(Gate g);

enum class TK : uint8_t { QUBIT, REC, SWEEP, PAULI, COMBINER };

struct Target {
    uint32_t value;   // qubit index, lookback k for rec[-k], sweep index
    TK kind;
    bool inverted;
    uint8_t pauli;    // bit0 = X component, bit1 = Z component (X=1, Z=2, Y=3) for PAULI targets
};

struct Instr {
    Gate gate;
    std::vector<double> args;
    std::vector<Target> targets;
    uint32_t tag = 0;   // marker set by host-side circuit rewrites (0 = none); the parser leaves it 0
};

struct Block;
struct Item {
    bool is_repeat = false;
    Instr instr;
    uint64_t reps = 0;
    std::shared_ptr<Block> body;
};
struct Block {
    std::vector<Item> items;
};

struct Stats {
    uint64_t num_qubits = 0;
    uint64_t num_measurements = 0;
    uint64_t num_detectors = 0;
    uint64_t num_observables = 0;
    uint64_t max_lookback = 0;
};

// Error kinds mapped to Python exceptions by the shim: invalid_argument -> ValueError,
// out_of_range -> IndexError (the reference's pybind11 default translation).
struct ParseError : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

Block parse_circuit(const char *text, size_t len);

// REPEAT-unrolled instruction stream (pointers into the block tree, which must outlive it).
std::vector<const Instr *> flatten(const Block &b);

Stats compute_stats(const std::vector<const Instr *> &flat);

// Number of measurement-record entries an instruction appends.
uint64_t num_results(const Instr &in);

// Pauli-product groups of an MPP/SPP instruction: [begin, end) index ranges into targets,
// skipping combiners.
std::vector<std::pair<size_t, size_t>> product_groups(const Instr &in);

}  // namespace sa
