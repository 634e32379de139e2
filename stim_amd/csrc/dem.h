// Detector error model (.dem) text: parsed and flattened into a list of error mechanisms.
//
// Follows the reference's file format (doc/file_format_dem_detector_error_model.md) and the sampler's
// reading of it (src/stim/simulators/dem_sampler.inl:50-77): every `error(p)` instruction, in the
// REPEAT-unrolled order, is one mechanism that flips its detector targets (D#, offset by the current
// `shift_detectors` total) and observable targets (L#); repeated targets cancel in pairs and `^`
// separators only suggest a decomposition. `detector` and `logical_observable` instructions declare
// indices without adding mechanisms.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sa {

struct DemError {
    double p;
    std::vector<uint32_t> dets;   // absolute detector indices (after cancellation, increasing)
    std::vector<uint32_t> obs;    // observable indices (after cancellation, increasing)
};

struct DemModel {
    uint64_t num_detectors = 0;
    uint64_t num_observables = 0;
    std::vector<DemError> errors;
};

// Throws std::invalid_argument with a description of the first problem.
DemModel parse_dem(const char *text, size_t len);

}  // namespace sa
