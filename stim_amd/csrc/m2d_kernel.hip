// Measurement-to-detection-event kernel for gfx950 (MI355X).
//
// One workgroup converts a tile of 32 shots. Its measurement rows are loaded (16-byte loads, several
// in flight per lane) and transposed 32x32 bits at a time across half-waves into measurement-major
// words in LDS (word m = measurement m of the tile's 32 shots). Each output column is then the XOR
// of its measurements' words, inverted when its reference parity is 1 (semantics in m2d.h), and the
// column words are transposed back into shot-major rows, staged in LDS when they fit and written as
// one contiguous block per tile. HBM-bound: each shot's measurement row is read once and its output
// rows written once.
#include <hip/hip_runtime.h>

#include "bitops.h"
#include "m2d.h"

using namespace sa;

namespace {

constexpr int NT = 512;
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

__device__ __forceinline__ uint32_t column_word(const M2dParams &P, const uint32_t *words, uint2 r) {
    const uint32_t c = (r.x >> 28) & 3u;
    uint32_t v;
    if (c == 3) {
        v = 0;
        const uint32_t n = P.ovf[r.y];
        for (uint32_t t = 0; t < n; t++) v ^= words[P.ovf[r.y + 1 + t]];
    } else {
        v = c >= 1 ? words[r.x & M2D_INDEX_MASK] : 0u;
        if (c == 2) v ^= words[r.y];
    }
    return (r.x & M2D_INVERT) ? ~v : v;
}

// Produces columns [c0, c0 + ncols) of the tile, transposed: lane l gets shot l's 4 bytes of each
// 32-column block, which go either to the LDS row image img (pitch bytes per shot) or straight to
// the shot's row in dst. CB column blocks per iteration keep that many record loads in flight.
__device__ __forceinline__ void emit_columns(const M2dParams &P, const uint32_t *words, uint8_t *img, uint32_t pitch,
                                             uint8_t *dst, uint64_t stride, uint32_t ncols, uint32_t c0,
                                             uint32_t half, uint32_t lane, bool have, uint64_t shot) {
    constexpr uint32_t CB = 4;
    const uint32_t nb = (ncols + 31) / 32;
    // Every lane of a half-wave runs the same iterations (the transpose shuffles within the half).
    for (uint32_t b0 = half * CB; b0 < nb; b0 += HALVES * CB) {
        uint2 r[CB];
#pragma unroll
        for (uint32_t u = 0; u < CB; u++) {
            const uint32_t c = (b0 + u) * 32 + lane;
            r[u] = c < ncols ? P.cols[c0 + c] : make_uint2(0, 0);
        }
#pragma unroll
        for (uint32_t u = 0; u < CB; u++) {
            const uint32_t b = b0 + u;
            const uint32_t v = transpose32(column_word(P, words, r[u]), lane);
            if (have && b < nb) {
                const uint32_t rem = ncols - b * 32;
                const uint32_t nbytes = rem >= 32 ? 4u : (rem + 7) / 8;
                if (img) {
                    uint8_t *q = img + lane * pitch + 4 * b;
                    for (uint32_t k = 0; k < nbytes; k++) q[k] = (uint8_t)(v >> (8 * k));
                } else {
                    store_bytes(dst + shot * stride + 4 * (uint64_t)b, v, nbytes);
                }
            }
        }
    }
}

// Copies the tile's staged rows (valid rows of pitch bytes) to dst rows [s0, s0 + valid): one
// contiguous block with 16-byte stores when the rows are packed back to back and aligned, else
// coalesced bytes.
__device__ __forceinline__ void flush_image(const uint8_t *img, uint32_t pitch, uint32_t valid, uint8_t *dst,
                                            uint64_t stride, uint64_t s0) {
    const uint32_t tid = threadIdx.x;
    uint8_t *base = dst + s0 * stride;
    const uint32_t total = valid * pitch;
    if (stride == pitch && ((uintptr_t)base & 15) == 0) {
        const uint32_t n16 = total / 16;
        for (uint32_t i = tid; i < n16; i += NT) {
            const uint4 v = *reinterpret_cast<const uint4 *>(img + 16 * i);
            __builtin_nontemporal_store(v.x, reinterpret_cast<uint32_t *>(base + 16 * i));
            __builtin_nontemporal_store(v.y, reinterpret_cast<uint32_t *>(base + 16 * i + 4));
            __builtin_nontemporal_store(v.z, reinterpret_cast<uint32_t *>(base + 16 * i + 8));
            __builtin_nontemporal_store(v.w, reinterpret_cast<uint32_t *>(base + 16 * i + 12));
        }
        for (uint32_t i = 16 * n16 + tid; i < total; i += NT) base[i] = img[i];
    } else {
        for (uint32_t i = tid; i < total; i += NT) {
            const uint32_t r = i / pitch, c = i % pitch;
            base[r * stride + c] = img[i];
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

    const uint32_t out_pitch = (P.n_out + 7) / 8, obs_pitch = (P.n_obs + 7) / 8;
    uint8_t *img_out = nullptr, *img_obs = nullptr;
    if (P.stage) {
        img_out = reinterpret_cast<uint8_t *>(words + (size_t)nblk * 32);
        img_obs = img_out + ((32 * out_pitch + 15) / 16) * 16;
    }
    emit_columns(P, words, img_out, out_pitch, P.out, P.out_stride, P.n_out, 0, half, lane, have, shot);
    if (P.n_obs)
        emit_columns(P, words, img_obs, obs_pitch, P.obs, P.obs_stride, P.n_obs, P.num_det, half, lane, have, shot);
    if (P.stage) {
        __syncthreads();
        if (P.n_out) flush_image(img_out, out_pitch, valid, P.out, P.out_stride, s0);
        if (P.n_obs) flush_image(img_obs, obs_pitch, valid, P.obs, P.obs_stride, s0);
    }
}

}  // namespace

namespace sa {

size_t m2d_lds_bytes(uint32_t blocks) {
    return (size_t)blocks * 32 * 4;
}

// LDS of a staged launch: the measurement words plus both 32-shot output images (16-byte aligned).
size_t m2d_staged_lds_bytes(const M2dParams &P) {
    return m2d_lds_bytes(P.blocks) + ((32 * (size_t)((P.n_out + 7) / 8) + 15) / 16) * 16 +
           ((32 * (size_t)((P.n_obs + 7) / 8) + 15) / 16) * 16;
}

hipError_t launch_m2d(const M2dParams &P, hipStream_t stream) {
    const uint64_t tiles = (P.shots + 31) / 32;
    if (tiles == 0) return hipSuccess;
    const size_t lds = P.stage ? m2d_staged_lds_bytes(P) : m2d_lds_bytes(P.blocks);
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
