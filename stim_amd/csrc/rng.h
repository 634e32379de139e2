// Counter and key construction of the kernels' Philox4x32-10 streams, shared by the device code and
// the host-side checks (plain C++; no device builtins).
//
// Every stream is keyed by the sampler's 64-bit key and numbered by a 64-bit global shot-tile index.
// A counter has four 32-bit words: (draw index, lane or item, low 32 bits of the tile, stream tag).
// The high 32 bits of the tile index go where the layout has room, so that a sampler's stream never
// repeats, and so that tiles below 2^32 keep exactly the counters (and outputs) they had before the
// high bits were carried:
//   - symptom kernels (scan and walk samplers): word 1 holds the lane (8 bits); the tile's high bits
//     sit above it. Exact (no two tiles share a counter) for tile indices below 2^56.
//   - frame kernel: every counter word is in use, so the high bits are XORed into the key's high
//     word. For one sampler the keys of different high parts are distinct; another sampler's key
//     (a 64-bit hash of its seed) matches one of them with probability 2^-64 per high part.
// The reference draws every sampler's randomness from one std::mt19937_64
// (src/stim/util_bot/probability_util.h:40-76), whose period is never reached in practice.
#pragma once

#include <cstdint>

#if defined(__HIP__)
#define SA_HD __host__ __device__
#else
#define SA_HD
#endif

namespace sa {

SA_HD inline uint32_t sym_lane_word(uint32_t lane, uint64_t tile) {
    return lane | ((uint32_t)(tile >> 32) << 8);
}

SA_HD inline uint32_t frame_key_hi(uint32_t k1, uint64_t tile) {
    return k1 ^ (uint32_t)(tile >> 32);
}

}  // namespace sa
