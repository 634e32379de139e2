// Measurement-to-detection-event kernel for gfx950 (MI355X).
//
// One workgroup converts a tile of 32 shots. Its measurement rows are loaded (16-byte loads, several
// in flight per lane) and transposed 32x32 bits at a time across half-waves into measurement-major
// words in LDS (word m = measurement m of the tile's 32 shots). Each output column is then the XOR
// of its measurements' words, inverted when its reference parity is 1 (semantics in m2d.h), and the
// column words are transposed back into shot-major rows, staged in LDS a chunk of columns at a time
// and copied to the rows with coalesced stores. HBM traffic is the minimum: each shot's measurement
// row is read once and its output rows written once.
//
// Work is laid out so that lane-dependent branches are rare: loads and LDS stores whose lanes fall
// past the tile's shots or the row's end go to clamped addresses or padding that no output reads.
#include <hip/hip_runtime.h>

#include "bitops.h"
#include "launch_util.h"
#include "m2d.h"

using namespace sa;

namespace {

constexpr int NT = 1024;
constexpr uint32_t HALVES = NT / 32;
constexpr uint32_t UNROLL = 4;                       // 16-byte loads in flight per lane
constexpr uint32_t CHUNK_BLOCKS = 64;                // output columns staged per pass: 64 blocks of 32
constexpr uint32_t CHUNK_BYTES = CHUNK_BLOCKS * 4;   // staged bytes per shot row
constexpr uint32_t PITCH = CHUNK_BYTES + 4;          // LDS bytes per staged row: the pad spreads the
                                                     // 32 rows of one column block over 32 banks

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
        const uint32_t n = r.x & M2D_INDEX_MASK;
        const uint4 *q = reinterpret_cast<const uint4 *>(P.ovf + r.y);
        for (uint32_t t = 0; t < n; t += 4) {
            const uint4 a = q[t / 4];
            v ^= words[a.x] ^ words[a.y] ^ words[a.z] ^ words[a.w];
        }
    } else {
        // Unused index fields hold 0, so both reads stay inside the tile's words.
        const uint32_t w0 = words[r.x & M2D_INDEX_MASK], w1 = words[r.y];
        v = (c >= 1 ? w0 : 0u) ^ (c == 2 ? w1 : 0u);
    }
    return (r.x & M2D_INVERT) ? ~v : v;
}

// Copies the staged chunk to the rows: staged row r's bytes [0, nbytes) go to bytes [off, off + nbytes)
// of shot row s0 + r, with coalesced byte stores (rows need no alignment). Thread t copies byte
// t % CHUNK_BYTES of rows t / CHUNK_BYTES + k * ROWS_PER_PASS: per pass the row base is the same for
// the whole workgroup and each thread's offset from it is fixed, so a store costs no address
// arithmetic. The offset (3 rows + 255 bytes) fits 32 bits because strides are below 2^30.
constexpr uint32_t ROWS_PER_PASS = NT / CHUNK_BYTES;
static_assert(NT % CHUNK_BYTES == 0 && 32 % ROWS_PER_PASS == 0, "flush passes cover whole tiles");

__device__ __forceinline__ void flush_chunk(const uint8_t *img, uint8_t *dst, uint64_t stride, uint32_t off,
                                            uint32_t nbytes, uint64_t s0, uint32_t valid) {
    const uint32_t c = threadIdx.x % CHUNK_BYTES, r0 = threadIdx.x / CHUNK_BYTES;
    if (c >= nbytes) return;
    const uint32_t voff = r0 * (uint32_t)stride + c;
    const uint8_t *const src = img + r0 * PITCH + c;
    uint8_t *const base = dst + s0 * stride + off;
    if (valid == 32) {
#pragma unroll
        for (uint32_t k = 0; k < 32 / ROWS_PER_PASS; k++) {
            __builtin_nontemporal_store(src[k * ROWS_PER_PASS * PITCH], base + k * ROWS_PER_PASS * stride + voff);
        }
    } else {
        for (uint32_t k = 0; r0 + k * ROWS_PER_PASS < valid; k++) {
            __builtin_nontemporal_store(src[k * ROWS_PER_PASS * PITCH], base + k * ROWS_PER_PASS * stride + voff);
        }
    }
}

// Produces columns [c0, c0 + ncols) of the tile (ncols >= 1), transposed, into shot rows dst (bit 0 =
// column c0), CHUNK_BLOCKS column blocks at a time: lane l of a half-wave gets shot l's 4 bytes of a
// 32-column block and puts them in the chunk's LDS image (PITCH bytes per shot); the chunk is then
// copied to the rows with coalesced byte stores (rows need no alignment). Image slots of lanes past
// the tile's shots or blocks past the last are written but never copied; columns past the last
// read as zero, so the padding bits of each row's last byte are 0.
__device__ __forceinline__ void emit_columns(const M2dParams &P, const uint32_t *words, uint8_t *img, uint8_t *dst,
                                             uint64_t stride, uint32_t ncols, uint32_t c0, uint32_t half,
                                             uint32_t lane, uint64_t s0, uint32_t valid) {
    constexpr uint32_t CB = CHUNK_BLOCKS / HALVES;
    const uint32_t nb = (ncols + 31) / 32, pitch = (ncols + 7) / 8;
    uint32_t *const slot = reinterpret_cast<uint32_t *>(img + lane * PITCH);
    for (uint32_t cb0 = 0; cb0 < nb; cb0 += CHUNK_BLOCKS) {
        // Every lane of a half-wave runs the same iterations (the transpose moves data within the half).
        uint2 r[CB];
#pragma unroll
        for (uint32_t u = 0; u < CB; u++) {
            const uint32_t c = (cb0 + half * CB + u) * 32 + lane;
            const uint2 rec = P.cols[c0 + (c < ncols ? c : ncols - 1)];
            r[u] = c < ncols ? rec : make_uint2(0, 0);
        }
#pragma unroll
        for (uint32_t u = 0; u < CB; u++) {
            slot[half * CB + u] = transpose32(column_word(P, words, r[u]), lane);
        }
        __syncthreads();
        const uint32_t off = 4 * cb0;
        const uint32_t nbytes = pitch - off < CHUNK_BYTES ? pitch - off : CHUNK_BYTES;
        flush_chunk(img, dst, stride, off, nbytes, s0, valid);
        __syncthreads();
    }
}

__global__ __launch_bounds__(NT) void m2d_kernel(M2dParams P) {
    extern __shared__ __align__(16) uint32_t words[];
    const uint32_t tid = threadIdx.x, half = tid >> 5, lane = tid & 31;
    const uint64_t s0 = (uint64_t)blockIdx.x * 32;
    const uint64_t left = P.shots - s0;
    const uint32_t valid = left < 32 ? (uint32_t)left : 32u;
    const bool have = lane < valid;
    // Lanes past the tile's shots read the tile's first row; their words are never used.
    const uint8_t *row = P.in + (s0 + (have ? lane : 0)) * P.in_stride;
    const uint32_t nblk = P.blocks;
    const uint32_t zero = m2d_zero_word(nblk, P.sweep_blocks);
    if (tid < 4) words[zero + tid] = 0;
    if (P.sweep_blocks) {
        // Sweep-bit words: each lane reads its shot's 4 bytes of a 32-bit block (zero without sweep
        // data), transposed like the measurement words.
        const uint8_t *srow = P.sweep ? P.sweep + (s0 + (have ? lane : 0)) * P.sweep_stride : nullptr;
        const uint32_t sw0 = m2d_sweep_word(nblk);
        for (uint32_t b = half; b < P.sweep_blocks; b += HALVES) {
            const uint32_t x = (have && srow) ? load_word_bytes(srow, P.sweep_row_bytes, b) : 0u;
            words[sw0 + b * 32 + lane] = transpose32(x, lane);
        }
    }

    if (P.in_vec) {
        // Groups of 4 blocks (16 bytes of each row); UNROLL groups per iteration. Loads past the last
        // group re-read it; the word region holds whole groups.
        const uint32_t ngrp = (nblk + 3) / 4;
        for (uint32_t g0 = half * UNROLL; g0 < ngrp; g0 += HALVES * UNROLL) {
            uint4 v[UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < UNROLL; u++) {
                const uint32_t g = g0 + u < ngrp ? g0 + u : ngrp - 1;
                v[u] = *reinterpret_cast<const uint4 *>(row + 16 * (uint64_t)g);
            }
#pragma unroll
            for (uint32_t u = 0; u < UNROLL; u++) {
                if (g0 + u < ngrp) {
                    const uint32_t xs[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                    for (uint32_t k = 0; k < 4; k++) words[(4 * (g0 + u) + k) * 32 + lane] = transpose32(xs[k], lane);
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

    uint8_t *const img = reinterpret_cast<uint8_t *>(words + zero + 4);
    if (P.n_out) emit_columns(P, words, img, P.out, P.out_stride, P.n_out, 0, half, lane, s0, valid);
    if (P.n_obs) emit_columns(P, words, img, P.obs, P.obs_stride, P.n_obs, P.num_det, half, lane, s0, valid);
}

// Random words of the free results: Philox4x32-10 keyed by the sampler's expansion key, counter (group of
// four free results, 64-bit tile index, stream tag).
constexpr uint32_t TAG_EXPAND = 7u << 28;

__global__ __launch_bounds__(NT) void expand_kernel(ExpandDev E) {
    extern __shared__ __align__(16) uint32_t words[];
    const M2dParams &P = E.out;
    const uint32_t tid = threadIdx.x, half = tid >> 5, lane = tid & 31;
    const uint64_t s0 = (uint64_t)blockIdx.x * 32;
    const uint64_t left = P.shots - s0;
    const uint32_t valid = left < 32 ? (uint32_t)left : 32u;
    const bool have = lane < valid;
    const uint32_t zero = m2d_zero_word(P.blocks, 0);
    for (uint32_t i = tid; i < zero + 4; i += NT) words[i] = 0;
    __syncthreads();
    // Free results (distinct from every contribution slot, so no ordering with the loads below).
    const uint64_t tile = E.tile0 + blockIdx.x;
    const uint32_t k0 = (uint32_t)E.seed, k1 = (uint32_t)(E.seed >> 32);
    for (uint32_t g = tid; g < (E.nfree + 3) / 4; g += NT) {
        const uint4 r = philox10(make_uint4(g, (uint32_t)tile, (uint32_t)(tile >> 32), TAG_EXPAND), k0, k1);
        const uint32_t rs[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            if (4 * g + k < E.nfree) words[E.free_slots[4 * g + k]] = rs[k];
        }
    }
    // y words, transposed across half-waves like the measurement words of m2d_kernel, XORed into their
    // slots. Lanes past the tile's shots read the tile's first row: those bits of every word belong to
    // shots that are never written.
    // Each lane's slot of its y column is loaded with the rows (it depends on nothing computed), so
    // a round's loads are all in flight together before the transposes and the XORs.
    const uint8_t *row = E.in + (s0 + (have ? lane : 0)) * E.in_stride;
    for (uint32_t g0 = half * UNROLL; g0 < E.in_groups; g0 += HALVES * UNROLL) {
        uint4 v[UNROLL];
        uint32_t pr[UNROLL][4];
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; u++) {
            const uint32_t g = g0 + u < E.in_groups ? g0 + u : E.in_groups - 1;
            v[u] = *reinterpret_cast<const uint4 *>(row + 16 * (uint64_t)g);
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) {
                const uint32_t y = (4 * (g0 + u) + k) * 32 + lane;
                pr[u][k] = (g0 + u < E.in_groups && y < E.Y) ? E.prim[y] : 0xFFFFFFFFu;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; u++) {
            if (g0 + u < E.in_groups) {
                const uint32_t xs[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (uint32_t k = 0; k < 4; k++) {
                    // Every lane of the half-wave takes part in the transpose.
                    const uint32_t x = transpose32(xs[k], lane);
                    if (pr[u][k] != 0xFFFFFFFFu) atomicXor(&words[pr[u][k] & 0xFFFFu], x);
                    if (pr[u][k] != 0xFFFFFFFFu && (pr[u][k] >> 31)) {
                        const uint32_t y = (4 * (g0 + u) + k) * 32 + lane;
                        for (uint32_t c = E.extra_off[y]; c < E.extra_off[y + 1]; c++) atomicXor(&words[E.extra[c]], x);
                    }
                }
            }
        }
    }
    __syncthreads();
    // Substitution: each thread runs its own items of a phase in record order (its chains need no
    // barrier); items read results of earlier phases or of the same thread.
    for (uint32_t p = 0; p < E.nphases; p++) {
        const uint32_t a = E.sched_off[p * NT + tid], b = E.sched_off[p * NT + tid + 1];
        for (uint32_t i = a; i < b; i++) {
            const uint4 r = E.items[i];
            uint32_t k = r.x & 0xFFFFu;
            const uint32_t n = r.x >> 16;
            uint32_t v = words[k];
            if (n == 1) {
                // A chain: each later result XORs the one before it, kept in v.
                v ^= words[r.y];
                words[k] = v;
                for (uint32_t j = 1; j < r.z; j++) {
                    k += r.w;
                    v ^= words[k];
                    words[k] = v;
                }
            } else {
                for (uint32_t t = 0; t < n; t++) v ^= words[E.terms[r.y + t]];
                words[k] = v;
            }
        }
        __syncthreads();
    }
    uint8_t *const img = reinterpret_cast<uint8_t *>(words + zero + 4);
    emit_columns(P, words, img, P.out, P.out_stride, P.n_out, 0, half, lane, s0, valid);
}

}  // namespace

namespace sa {

hipError_t launch_expand(const ExpandDev &E, hipStream_t stream) {
    const uint64_t tiles = (E.out.shots + 31) / 32;
    if (tiles == 0) return hipSuccess;
    const size_t lds = m2d_lds_bytes(E.out.blocks, 0);
    static size_t reserved = 0;
    const hipError_t e = reserve_lds(reinterpret_cast<const void *>(&expand_kernel), lds, reserved);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(expand_kernel, dim3((uint32_t)tiles), dim3(NT), lds, stream, E);
    return hipGetLastError();
}

// LDS of a launch: the tile's measurement, sweep and zero words plus the staged output chunk.
size_t m2d_lds_bytes(uint32_t blocks, uint32_t sweep_blocks) {
    return ((size_t)m2d_zero_word(blocks, sweep_blocks) + 4) * 4 + 32 * PITCH;
}

hipError_t launch_m2d(const M2dParams &P, hipStream_t stream) {
    const uint64_t tiles = (P.shots + 31) / 32;
    if (tiles == 0) return hipSuccess;
    const size_t lds = m2d_lds_bytes(P.blocks, P.sweep_blocks);
    static size_t reserved = 0;
    const hipError_t e = reserve_lds(reinterpret_cast<const void *>(&m2d_kernel), lds, reserved);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(m2d_kernel, dim3((uint32_t)tiles), dim3(NT), lds, stream, P);
    return hipGetLastError();
}

}  // namespace sa
