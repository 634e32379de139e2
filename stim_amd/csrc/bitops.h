// Device helpers shared by the kernels (gfx950, 64-lane waves): the Philox4x32-10 generator and
// bit-matrix helpers (the transpose works on the 32 lanes of a half-wave).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sa {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11).
// Each round's products are formed as one 32x32->64 multiply per multiplier (v_mad_u64_u32) rather
// than separate high and low halves. Counters here mix words that are uniform across a wave (draw
// index, tile, stream tag) with a per-lane word; the first rounds are left to the compiler, which
// keeps the still-uniform words on the scalar unit. By round 4 every word differs per lane, and from
// there each output word's two XORs (with the other word and the round key) are one three-input
// v_bitop3_b32 (truth table 0x96 = XOR3).
__device__ __forceinline__ uint4 philox10(uint4 c, uint32_t k0, uint32_t k1) {
    // The keys are uniform across the wave. Hiding their values here makes the compiler form the
    // round keys per call with scalar adds, instead of hoisting all 20 out of the caller's loops,
    // where they outlive the scalar registers and spill into vector-register lanes.
    asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
#pragma unroll
    for (int r = 4; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4(__builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c.y, k0, 0x96), (uint32_t)p1,
                       __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c.w, k1, 0x96), (uint32_t)p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// x of lane (lane ^ J) within each 32-lane half of the wave. Partners 1, 2 and 8 apart are DPP moves
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror:8 within 16-lane rows); 4 and 16 apart use
// ds_swizzle's bit-mask mode (and 0x1F, xor J), which permutes within 32-lane groups.
template <int J>
__device__ __forceinline__ uint32_t xor_lanes(uint32_t x) {
    if constexpr (J == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true);
    } else if constexpr (J == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true);
    } else if constexpr (J == 8) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, true);
    } else {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (J << 10) | 0x1F);
    }
}

// One level of the transpose. With m the level's low-half mask (0x0000FFFF for J = 16 down to
// 0x55555555 for J = 1): lanes without bit J keep their low halves and take the partner's low
// halves shifted up by J; lanes with bit J keep their high halves and take the partner's high
// halves shifted down. Both moves are one rotate of the partner's word (right by J, or by 32 - J),
// merged with one bit-field select: keep = m or ~m, rot = 32 - J or J, per lane.
// Bit-field select (keep & a) | (~keep & b) as one v_bfi_b32: written as the plain expression, the
// compiler hoists keep and ~keep as separate loop invariants and emits an AND plus an AND-OR.
__device__ __forceinline__ uint32_t bit_select(uint32_t keep, uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(keep), "v"(a), "v"(b));
    return r;
}

template <int J>
__device__ __forceinline__ uint32_t transpose_level(uint32_t x, uint32_t keep, uint32_t rot) {
    const uint32_t y = xor_lanes<J>(x);
    const uint32_t ys = __builtin_amdgcn_alignbit(y, y, rot);
    return bit_select(keep, x, ys);
}

template <int J>
__device__ __forceinline__ uint32_t transpose_level(uint32_t x, uint32_t m, uint32_t lane, int) {
    const bool hi = lane & J;
    return transpose_level<J>(x, hi ? ~m : m, hi ? J : 32 - J);
}

// 32x32 bit transpose across the 32 lanes of a half-wave: lane r holds row r on input and
// column r on output (recursive block swap, one cross-lane move per level). The per-lane masks and
// rotate counts depend only on the lane, so loops keep them in registers.
__device__ __forceinline__ uint32_t transpose32(uint32_t x, uint32_t lane) {
    x = transpose_level<16>(x, 0x0000FFFFu, lane, 0);
    x = transpose_level<8>(x, 0x00FF00FFu, lane, 0);
    x = transpose_level<4>(x, 0x0F0F0F0Fu, lane, 0);
    x = transpose_level<2>(x, 0x33333333u, lane, 0);
    x = transpose_level<1>(x, 0x55555555u, lane, 0);
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
