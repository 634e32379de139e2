// Reference samples: the noiseless run of a circuit on a stabilizer tableau, with every random
// measurement result taken as 0.
//
// Reference: TableauSimulator::reference_sample_circuit (src/stim/simulators/tableau_simulator.inl:1435-1438)
// runs the circuit's noiseless form (Circuit::aliased_noiseless_circuit, src/stim/circuit/circuit.cc:775-807:
// noise dropped, result-flip probabilities dropped, heralded errors recorded as 0) with sign bias +1,
// so a random measurement collapses onto the +1 eigenstate of the measured observable (result 0
// before any inversion of the recorded bit). Measurement samplers XOR this sample into the frame
// simulator's flips (src/stim/simulators/frame_simulator_util.inl:290-372).
#pragma once

#include <cstdint>
#include <vector>

#include "circuit.h"

namespace sa {

// Random collapses of a reference run: those that produced a recorded result, and all of them (a
// reset of a qubit whose result was random also collapses). When every random collapse is recorded,
// the noiseless record is uniform over an affine space of dimension random_recorded.
struct RefStats {
    uint64_t random_recorded = 0;
    uint64_t random_total = 0;
};

// One byte (0 or 1) per measurement result, in record order.
std::vector<uint8_t> reference_sample(const std::vector<const Instr *> &flat, RefStats *stats = nullptr);

}  // namespace sa
