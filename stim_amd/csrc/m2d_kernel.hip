// Measurement-to-detection-event kernel for gfx950 (MI355X).
//
// One workgroup converts a tile of 32 shots. Its measurement rows are loaded (16-byte loads, several
// in flight per lane) and transposed 32x32 bits at a time across half-waves into measurement-major
// words in LDS (word m = measurement m of the tile's 32 shots). Each output column is then the XOR
// of its measurements' words, inverted when its reference parity is 1 (semantics in m2d.h), and the
// column words are transposed back into the shot-major output rows. HBM-bound: each shot's
// measurement row is read once and its output rows written once.
#include <hip/hip_runtime.h>

#include "bitops.h"
#include "m2d.h"

using namespace sa;

namespace {

constexpr int NT = 256;
constexpr uint32_t HALVES = NT / 32;
constexpr uint32_t UNROLL = 4;   // 16-byte loads in flight per lane

// Bytes [4b, 4b + 4) of a row, zero past the row's end.
__device__ __forceinline__ uint32_t load_word_bytes(const uint8_t *row, uint32_t row_bytes, uint32_t b) {
    uint32_t x = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        const uint32_t i = 4 * b + k;
        if (i < row_bytes) x |= (uint32_t)row[i] << (8 * k);
    }
    return x;
}

// Writes columns [c0, c0 + ncols) of the tile, transposed, into shot rows dst (bit 0 = column c0).
__device__ __forceinline__ void emit_columns(const M2dParams &P, const uint32_t *words, uint8_t *dst, uint64_t stride,
                                             uint32_t ncols, uint32_t c0, uint32_t half, uint32_t lane, bool have,
                                             uint64_t shot) {
    const uint32_t nb = (ncols + 31) / 32;
    // Every lane of a half-wave runs the same iterations (the transpose shuffles within the half).
    for (uint32_t b = half; b < nb; b += HALVES) {
        const uint32_t c = b * 32 + lane;
        uint32_t v = 0;
        if (c < ncols) {
            const uint32_t col = c0 + c;
            const uint32_t e = P.off[col + 1];
            for (uint32_t t = P.off[col]; t < e; t++) v ^= words[P.terms[t]];
            if ((P.flip[col >> 5] >> (col & 31)) & 1u) v = ~v;
        }
        v = transpose32(v, lane);
        if (have) {
            const uint32_t rem = ncols - b * 32;
            store_bytes(dst + shot * stride + 4 * (uint64_t)b, v, rem >= 32 ? 4u : (rem + 7) / 8);
        }
    }
}

__global__ __launch_bounds__(NT) void m2d_kernel(M2dParams P) {
    extern __shared__ __align__(16) uint32_t words[];
    const uint32_t tid = threadIdx.x, half = tid >> 5, lane = tid & 31;
    const uint64_t s0 = (uint64_t)blockIdx.x * 32;
    const uint64_t left = P.shots - s0;
    const uint32_t valid = left < 32 ? (uint32_t)left : 32u;
    const bool have = lane < valid;
    const uint64_t shot = s0 + lane;
    const uint8_t *row = P.in + (have ? shot : s0) * P.in_stride;
    const uint32_t nblk = P.blocks;

    if (P.in_vec) {
        // Groups of 4 blocks (16 bytes of each row); UNROLL groups per iteration.
        const uint32_t ngrp = (nblk + 3) / 4;
        for (uint32_t g0 = half * UNROLL; g0 < ngrp; g0 += HALVES * UNROLL) {
            uint4 v[UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < UNROLL; u++) {
                const uint32_t g = g0 + u;
                v[u] = (have && g < ngrp) ? *reinterpret_cast<const uint4 *>(row + 16 * (uint64_t)g) : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (uint32_t u = 0; u < UNROLL; u++) {
                const uint32_t xs[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (uint32_t k = 0; k < 4; k++) {
                    const uint32_t x = transpose32(xs[k], lane);
                    const uint32_t blk = 4 * (g0 + u) + k;
                    if (blk < nblk) words[blk * 32 + lane] = x;
                }
            }
        }
    } else {
        for (uint32_t b = half; b < nblk; b += HALVES) {
            const uint32_t x = transpose32(have ? load_word_bytes(row, P.in_row_bytes, b) : 0u, lane);
            words[b * 32 + lane] = x;
        }
    }
    __syncthreads();

    emit_columns(P, words, P.out, P.out_stride, P.n_out, 0, half, lane, have, shot);
    if (P.n_obs) emit_columns(P, words, P.obs, P.obs_stride, P.n_obs, P.num_det, half, lane, have, shot);
}

}  // namespace

namespace sa {

size_t m2d_lds_bytes(uint32_t blocks) {
    return (size_t)blocks * 32 * 4;
}

hipError_t launch_m2d(const M2dParams &P, hipStream_t stream) {
    const uint64_t tiles = (P.shots + 31) / 32;
    if (tiles == 0) return hipSuccess;
    const size_t lds = m2d_lds_bytes(P.blocks);
    static size_t configured = 0;
    if (lds > configured) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&m2d_kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        configured = lds;
    }
    hipLaunchKernelGGL(m2d_kernel, dim3((uint32_t)tiles), dim3(NT), lds, stream, P);
    return hipGetLastError();
}

}  // namespace sa
