// Bit-matrix helpers shared by the device kernels (gfx950, 64-lane waves; the transpose works on
// the 32 lanes of a half-wave).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sa {

// 32x32 bit transpose across the 32 lanes of a half-wave: lane r holds row r on input and
// column r on output (recursive block swap, one shuffle per level).
__device__ __forceinline__ uint32_t transpose32(uint32_t x, uint32_t lane) {
    const uint32_t masks[5] = {0x0000FFFFu, 0x00FF00FFu, 0x0F0F0F0Fu, 0x33333333u, 0x55555555u};
#pragma unroll
    for (int lvl = 0; lvl < 5; lvl++) {
        const int j = 16 >> lvl;
        const uint32_t m = masks[lvl];
        uint32_t y = __shfl_xor(x, j, 32);
        if (lane & j) {
            x = (x & ~m) | ((y >> j) & m);
        } else {
            x = (x & m) | ((y & m) << j);
        }
    }
    return x;
}

// Streams up to 4 bytes of x to p with the widest aligned stores (non-temporal: output rows are
// written once and must not evict the program from L2).
__device__ __forceinline__ void store_bytes(uint8_t *p, uint32_t x, uint32_t nbytes) {
    const uintptr_t a = (uintptr_t)p;
    if (nbytes == 4 && (a & 3) == 0) {
        __builtin_nontemporal_store(x, reinterpret_cast<uint32_t *>(p));
    } else if (nbytes == 4 && (a & 1) == 0) {
        __builtin_nontemporal_store((uint16_t)x, reinterpret_cast<uint16_t *>(p));
        __builtin_nontemporal_store((uint16_t)(x >> 16), reinterpret_cast<uint16_t *>(p + 2));
    } else {
        for (uint32_t k = 0; k < nbytes; k++) __builtin_nontemporal_store((uint8_t)(x >> (8 * k)), p + k);
    }
}

}  // namespace sa
