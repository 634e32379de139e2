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
//
// Also here: the windowed conversion kernel (a ring of recent measurement words per tile, for records
// longer than LDS holds; m2d.h) and the measurement sampler's expansion kernel (detection events to
// measurement records; expand.h).
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

// Word i of the tile's words, read at its LDS byte address (the words start at LDS address `base`,
// 0 in these kernels, so no base is added per read).
typedef __attribute__((address_space(3))) const uint32_t lds_cu32;
__device__ __forceinline__ uint32_t lds_word(uint32_t base, uint32_t i) {
    return *(lds_cu32 *)(uintptr_t)(base + 4 * i);
}

__device__ __forceinline__ uint32_t column_word(const M2dParams &P, uint32_t base, uint2 r) {
    uint32_t v;
    if (((r.x >> 28) & 3u) == 3u) {
        v = 0;
        const uint32_t n = r.x & M2D_INDEX_MASK;
        const uint4 *q = reinterpret_cast<const uint4 *>(P.ovf + r.y);
        for (uint32_t t = 0; t < n; t += 4) {
            const uint4 a = q[t / 4];
            v ^= lds_word(base, a.x) ^ lds_word(base, a.y) ^ lds_word(base, a.z) ^ lds_word(base, a.w);
        }
    } else {
        // Unused index fields name the zero word, so both words are XORed unconditionally.
        v = lds_word(base, r.x & M2D_INDEX_MASK) ^ lds_word(base, r.y);
    }
    return v ^ (0u - (r.x >> 31));
}

// Copies staged rows [0, valid) (pitch P bytes in LDS) to bytes [off, off + nbytes) of shot rows s0 + r
// as 4-byte stores at any byte alignment (the HSA runtime enables unaligned global access; the
// compiler emits global_store_dword for an align-1 word): lane w of a wave takes bytes [4w, 4w + 4),
// so one store instruction writes 256 contiguous bytes of a row; the wave's rows step by NTHREADS /
// 64. A tail of 1 to 3 bytes at the end of a row is written byte by byte.
typedef uint32_t unaligned_u32 __attribute__((aligned(1)));
template <uint32_t NTHREADS, uint32_t P>
__device__ __forceinline__ void flush_words(const uint8_t *img, uint8_t *dst, uint64_t stride, uint32_t off,
                                            uint32_t nbytes, uint64_t s0, uint32_t valid) {
    constexpr uint32_t RG = NTHREADS / 64;
    static_assert(P % 4 == 0 && NTHREADS % 64 == 0, "word-aligned staging rows, whole waves");
    const uint32_t b = 4 * (threadIdx.x % 64), r0 = threadIdx.x / 64;
    if (b >= nbytes) return;
    uint8_t *const base = dst + s0 * stride + off + b;
    const uint8_t *const src = img + b;
    if (b + 4 <= nbytes) {
        for (uint32_t r = r0; r < valid; r += RG) {
            const uint32_t x = *reinterpret_cast<const uint32_t *>(src + r * P);
            __builtin_nontemporal_store(x, reinterpret_cast<unaligned_u32 *>(base + r * stride));
        }
    } else {
        for (uint32_t r = r0; r < valid; r += RG) {
            for (uint32_t k = 0; b + k < nbytes; k++) __builtin_nontemporal_store(src[r * P + k], base + r * stride + k);
        }
    }
}

// Produces columns [c0, c0 + ncols) of the tile (ncols >= 1), transposed, into shot rows dst (bit 0 =
// column c0), CHUNK_BLOCKS column blocks at a time: lane l of a half-wave gets shot l's 4 bytes of a
// 32-column block and puts them in the chunk's LDS image (PITCH bytes per shot); the chunk is then
// copied to the rows with 4-byte stores at any alignment (flush_words). Image slots of lanes past
// the tile's shots or blocks past the last are written but never copied; columns past the last
// read as zero, so the padding bits of each row's last byte are 0.
__device__ __forceinline__ void emit_columns(const M2dParams &P, const uint32_t *words, uint8_t *img, uint8_t *dst,
                                             uint64_t stride, uint32_t ncols, uint32_t c0, uint32_t half,
                                             uint32_t lane, uint64_t s0, uint32_t valid) {
    constexpr uint32_t CB = CHUNK_BLOCKS / HALVES;
    const uint32_t nb = (ncols + 31) / 32, pitch = (ncols + 7) / 8;
    uint32_t *const slot = reinterpret_cast<uint32_t *>(img + lane * PITCH);
    const uint32_t base = (uint32_t)(uintptr_t)(lds_cu32 *)words;
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
            slot[half * CB + u] = transpose32(column_word(P, base, r[u]), lane);
        }
        __syncthreads();
        const uint32_t off = 4 * cb0;
        const uint32_t nbytes = pitch - off < CHUNK_BYTES ? pitch - off : CHUNK_BYTES;
        flush_words<NT, PITCH>(img, dst, stride, off, nbytes, s0, valid);
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

// The expansion's output: column m of the rows is measurement word m, inverted where the reference
// sample is 1. Column inversion flips bit m of every shot's row, which after the transpose is one XOR
// of the block's 32 inversion bits (inv[b], the same for every lane) into each lane's output word, so
// no column records are read. Words past the last measurement are zero, as are their inversion bits.
__device__ __forceinline__ void emit_identity(const uint32_t *words, const uint32_t *inv, uint8_t *img, uint8_t *dst,
                                              uint64_t stride, uint32_t ncols, uint32_t half, uint32_t lane,
                                              uint64_t s0, uint32_t valid) {
    constexpr uint32_t CB = CHUNK_BLOCKS / HALVES;
    const uint32_t nb = (ncols + 31) / 32, pitch = (ncols + 7) / 8;
    uint32_t *const slot = reinterpret_cast<uint32_t *>(img + lane * PITCH);
    for (uint32_t cb0 = 0; cb0 < nb; cb0 += CHUNK_BLOCKS) {
        uint32_t w[CB], iv[CB];
#pragma unroll
        for (uint32_t u = 0; u < CB; u++) {
            const uint32_t b = cb0 + half * CB + u;
            const uint32_t bc = b < nb ? b : nb - 1;
            w[u] = words[bc * 32 + lane];
            iv[u] = inv[bc];
        }
#pragma unroll
        for (uint32_t u = 0; u < CB; u++) slot[half * CB + u] = transpose32(w[u], lane) ^ iv[u];
        __syncthreads();
        const uint32_t off = 4 * cb0;
        const uint32_t nbytes = pitch - off < CHUNK_BYTES ? pitch - off : CHUNK_BYTES;
        flush_words<NT, PITCH>(img, dst, stride, off, nbytes, s0, valid);
        __syncthreads();
    }
}

// Random words of the free results: Philox4x32-10 keyed by the sampler's expansion key, counter (group of
// four free results, 64-bit tile index, stream tag).
constexpr uint32_t TAG_EXPAND = 7u << 28;

// One substitution record (layout in expand.h): words r.x .. r.w, q.x .. q.w.
__device__ __forceinline__ void expand_record(const ExpandDev &E, uint32_t *words, uint4 r, uint4 q) {
    uint32_t k = r.x & 0xFFFFu;
    const uint32_t n = r.x >> 16;
    uint32_t v = words[k];
    if (n == 1) {
        // A chain: each later result XORs the one before it, kept in v. Its results are distinct
        // (stride w > 0) and only this record writes them, so the words it reads are the values
        // before the chain: they are read 8 at a time ahead of the writes, which turns the chain's
        // dependent LDS round trips into one per 8 steps.
        v ^= words[r.y];
        words[k] = v;
        uint32_t j = 1;
        for (; j + 8 <= r.z; j += 8) {
            uint32_t t[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; u++) t[u] = words[k + (u + 1) * r.w];
#pragma unroll
            for (uint32_t u = 0; u < 8; u++) {
                v ^= t[u];
                words[k + (u + 1) * r.w] = v;
            }
            k += 8 * r.w;
        }
        for (; j < r.z; j++) {
            k += r.w;
            v ^= words[k];
            words[k] = v;
        }
    } else if (n <= EXPAND_INLINE_TERMS) {
        // Terms inline: all seven read together (unused ones name result 0 and are masked).
        const uint32_t t[EXPAND_INLINE_TERMS] = {r.y, r.z, r.w, q.x, q.y, q.z, q.w};
        uint32_t w[EXPAND_INLINE_TERMS];
#pragma unroll
        for (uint32_t u = 0; u < EXPAND_INLINE_TERMS; u++) w[u] = words[t[u]];
#pragma unroll
        for (uint32_t u = 0; u < EXPAND_INLINE_TERMS; u++) v ^= u < n ? w[u] : 0u;
        words[k] = v;
    } else {
        for (uint32_t t = 0; t < n; t++) v ^= words[E.terms[r.y + t]];
        words[k] = v;
    }
}

// Two workgroups per CU (their LDS words allow two): 8 waves per SIMD, at most 64 VGPRs.
__global__ __launch_bounds__(NT, 8) void expand_kernel(ExpandDev E) {
    extern __shared__ __align__(16) uint32_t words[];
    const M2dParams &P = E.out;
    const uint32_t tid = threadIdx.x, half = tid >> 5, lane = tid & 31;
    const uint64_t s0 = (uint64_t)blockIdx.x * 32;
    const uint64_t left = P.shots - s0;
    const uint32_t valid = left < 32 ? (uint32_t)left : 32u;
    const bool have = lane < valid;
    const uint32_t zero = m2d_zero_word(P.blocks, 0);
    // Phase 0's record range, loaded ahead (phase p + 1's is loaded while phase p runs).
    // (sched_off holds nphases * NT + 1 offsets: no range without a phase.)
    uint32_t na = E.nphases ? E.sched_off[tid] : 0u, nb = E.nphases ? E.sched_off[tid + 1] : 0u;
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
    // Phase 0's first record (record 0 for an empty range; items holds at least one), loaded while
    // the other waves finish the y pass.
    uint4 nr = E.items[2 * (na < nb ? na : 0)], nq = E.items[2 * (na < nb ? na : 0) + 1];
    __syncthreads();
    // Substitution: each thread runs its own items of a phase in record order (its chains need no
    // barrier); items read results of earlier phases or of the same thread. Each phase's record range
    // is loaded while the previous phase runs and its first record before the previous barrier.
    for (uint32_t p = 0; p < E.nphases; p++) {
        const uint32_t a = na, b = nb;
        uint4 r = nr, q = nq;
        if (p + 1 < E.nphases) {
            na = E.sched_off[(p + 1) * NT + tid];
            nb = E.sched_off[(p + 1) * NT + tid + 1];
        }
        // A thread's records of a phase may depend on each other (through results in LDS only): the
        // next record's load is issued before the current one runs.
        for (uint32_t i = a; i < b; i++) {
            const uint32_t nx = i + 1 < b ? i + 1 : i;
            const uint4 r2 = E.items[2 * nx], q2 = E.items[2 * nx + 1];
            expand_record(E, words, r, q);
            r = r2;
            q = q2;
        }
        if (p + 1 < E.nphases) {
            const uint32_t f = na < nb ? na : 0;
            nr = E.items[2 * f];
            nq = E.items[2 * f + 1];
        }
        __syncthreads();
    }
    uint8_t *const img = reinterpret_cast<uint8_t *>(words + zero + 4);
    emit_identity(words, E.inv, img, P.out, P.out_stride, P.n_out, half, lane, s0, valid);
}

// ------------------------------------------------------------------ windowed conversion
constexpr int NTW = 256;
constexpr uint32_t WIN_HALVES = NTW / 32;
constexpr uint32_t WIN_STAGE_BLOCKS = 64;                 // output blocks staged per flush (256 bytes per row)
constexpr uint32_t WIN_PITCH = WIN_STAGE_BLOCKS * 4 + 4;  // LDS bytes per staged row

// A column word from the ring: term m is ring word m & mask; overflow padding (an index >= M) is 0.
__device__ __forceinline__ uint32_t ring_column_word(const M2dParams &P, const uint32_t *ring, uint32_t mask,
                                                     uint32_t M, uint2 r) {
    const uint32_t c = (r.x >> 28) & 3u;
    uint32_t v;
    if (c == 3) {
        v = 0;
        const uint32_t n = r.x & M2D_INDEX_MASK;
        const uint4 *q = reinterpret_cast<const uint4 *>(P.ovf + r.y);
        for (uint32_t t = 0; t < n; t += 4) {
            const uint4 a = q[t / 4];
            v ^= (a.x < M ? ring[a.x & mask] : 0u) ^ (a.y < M ? ring[a.y & mask] : 0u) ^
                 (a.z < M ? ring[a.z & mask] : 0u) ^ (a.w < M ? ring[a.w & mask] : 0u);
        }
    } else {
        const uint32_t w0 = ring[(r.x & M2D_INDEX_MASK) & mask], w1 = ring[r.y & mask];
        v = (c >= 1 ? w0 : 0u) ^ (c == 2 ? w1 : 0u);
    }
    return (r.x & M2D_INVERT) ? ~v : v;
}

// Writes staged blocks [0, nb) (bytes [0, nbytes) of each staged row) to bytes [off, off + nbytes) of
// the tile's rows: thread t copies byte t of every row (coalesced across the workgroup per row).
__device__ __forceinline__ void win_flush(const uint8_t *img, uint8_t *dst, uint64_t stride, uint32_t off,
                                          uint32_t nbytes, uint64_t s0, uint32_t valid) {
    const uint32_t c = threadIdx.x;
    if (c >= nbytes) return;
    uint8_t *base = dst + s0 * stride + off + c;
    for (uint32_t r = 0; r < valid; r++) __builtin_nontemporal_store(img[r * WIN_PITCH + c], base + r * stride);
}

// Output blocks are staged a pass at a time: up to WIN_PASS blocks, block b + u * WIN_HALVES + half
// on half-wave `half` for u < WIN_UNROLL, each lane producing one column of its block.
constexpr uint32_t WIN_UNROLL = 4;
constexpr uint32_t WIN_PASS = WIN_HALVES * WIN_UNROLL;

// Column records of a pass of n blocks from block b; lanes past n or past the part's columns load the
// zero record (an address select, so the loads stay in flight until the records are used).
__device__ __forceinline__ void win_records(const M2dParams &P, uint2 (&r)[WIN_UNROLL], uint32_t ncols,
                                            uint32_t col0, uint32_t zero_rec, uint32_t b, uint32_t n) {
    const uint32_t half = threadIdx.x >> 5, lane = threadIdx.x & 31;
#pragma unroll
    for (uint32_t u = 0; u < WIN_UNROLL; u++) {
        const uint32_t i = u * WIN_HALVES + half;
        const uint32_t c = (b + i) * 32 + lane;
        r[u] = P.cols[(i < n && c < ncols) ? col0 + c : zero_rec];
    }
}

// Emission state of one part: `staged` blocks in the staging image since the last flush, `flushed`
// blocks already written to the rows.
struct WinPart {
    uint8_t *dst;
    uint64_t stride;
    uint32_t ncols, col0, flushed, staged;
};

// Blocks of the next pass from block b of [b, b1): bounded by the staging room and WIN_PASS.
__device__ __forceinline__ uint32_t win_pass_blocks(const WinPart &W, uint32_t b, uint32_t b1) {
    const uint32_t room = WIN_STAGE_BLOCKS - W.staged, left = b1 - b;
    const uint32_t n = left < room ? left : room;
    return n < WIN_PASS ? n : WIN_PASS;
}

// Emits blocks [b, b1) of a part, the first pass from the records r (loaded by the caller for the
// pass win_pass_blocks gives), and flushes the staging image whenever it fills and, with final_flush,
// at the end. Uniform across the workgroup.
__device__ __forceinline__ void win_emit(const M2dParams &P, const uint32_t *ring, uint32_t mask, uint32_t M,
                                         uint32_t zero_rec, uint8_t *img, WinPart &W, uint32_t b, uint32_t b1,
                                         uint2 (&r)[WIN_UNROLL], uint64_t s0, uint32_t valid, bool final_flush) {
    const uint32_t half = threadIdx.x >> 5, lane = threadIdx.x & 31;
    const uint32_t pitch = (W.ncols + 7) / 8;
    uint32_t *const slot = reinterpret_cast<uint32_t *>(img + lane * WIN_PITCH);
    bool first = true;
    while (b < b1 || (final_flush && W.staged)) {
        const uint32_t n = win_pass_blocks(W, b, b1);
        if (!first) win_records(P, r, W.ncols, W.col0, zero_rec, b, n);
        first = false;
        // Every lane of a half-wave takes part in the transposes; the zero record yields 0.
#pragma unroll
        for (uint32_t u = 0; u < WIN_UNROLL; u++) {
            const uint32_t i = u * WIN_HALVES + half;
            const uint32_t x = transpose32(ring_column_word(P, ring, mask, M, r[u]), lane);
            if (i < n) slot[W.staged + i] = x;
        }
        b += n;
        W.staged += n;
        if (W.staged == WIN_STAGE_BLOCKS || (final_flush && b == b1)) {
            __syncthreads();
            const uint32_t off = 4 * W.flushed;
            const uint32_t nbytes = pitch - off < 4 * W.staged ? pitch - off : 4 * W.staged;
            flush_words<NTW, WIN_PITCH>(img, W.dst, W.stride, off, nbytes, s0, valid);
            __syncthreads();
            W.flushed += W.staged;
            W.staged = 0;
        }
    }
}

// Windowed conversion of a 32-shot tile (semantics in m2d.h). Per chunk: its 16-byte groups (one per
// half-wave, loaded during the previous chunk) go into the ring; the first pass's column records are
// loaded, then the next chunk's groups (loads complete in order, so waiting for the records leaves
// the row load in flight); then the chunk's ready blocks are emitted.
__global__ __launch_bounds__(NTW) void m2d_win_kernel(M2dParams P, M2dWindow Wn) {
    extern __shared__ __align__(16) uint32_t ring[];
    const uint32_t tid = threadIdx.x, half = tid >> 5, lane = tid & 31;
    const uint64_t s0 = (uint64_t)blockIdx.x * 32;
    const uint64_t left = P.shots - s0;
    const uint32_t valid = left < 32 ? (uint32_t)left : 32u;
    const bool have = lane < valid;
    // Lanes past the tile's shots read the tile's first row; their words are never used.
    const uint4 *row = reinterpret_cast<const uint4 *>(P.in + (s0 + (have ? lane : 0)) * P.in_stride);
    const uint32_t mask = Wn.ring_words - 1, M = P.blocks * 32;
    const uint32_t ngrp = (P.blocks + 3) / 4;
    uint8_t *const img = reinterpret_cast<uint8_t *>(ring + Wn.ring_words);
    uint32_t *const chunk_end = reinterpret_cast<uint32_t *>(img + 32 * WIN_PITCH);
    for (uint32_t i = tid; i < Wn.nchunks; i += NTW) chunk_end[i] = Wn.chunk_end[i];
    const uint32_t nb0 = (P.n_out + 31) / 32;
    constexpr uint32_t GPC = M2D_WIN_CHUNK / 128;   // 16-byte groups per chunk, one per half-wave
    static_assert(GPC == WIN_HALVES, "one group per half-wave and chunk");
    // Loads past the row's last group re-read it (their words are not stored).
    uint4 v = row[half < ngrp ? half : ngrp - 1];
    WinPart W0{P.out, P.out_stride, P.n_out, 0, 0, 0};
    uint32_t next0 = 0;
    uint2 r[WIN_UNROLL];
    for (uint32_t ch = 0; ch < Wn.nchunks; ch++) {
        const uint32_t g = ch * GPC + half;
        __syncthreads();   // the previous chunk's columns are read before its ring slots are reused
        {
            const uint32_t xs[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) {
                const uint32_t x = transpose32(xs[k], lane);
                if (g < ngrp) ring[((4 * g + k) * 32 + lane) & mask] = x;
            }
        }
        const uint32_t b1 = chunk_end[ch];
        win_records(P, r, W0.ncols, 0, Wn.zero_rec, next0, win_pass_blocks(W0, next0, b1));
        const uint32_t gn = g + GPC;
        v = row[gn < ngrp ? gn : ngrp - 1];
        __syncthreads();
        if (b1 > next0) win_emit(P, ring, mask, M, Wn.zero_rec, img, W0, next0, b1, r, s0, valid, false);
        next0 = b1;
    }
    if (P.n_out) {
        win_records(P, r, W0.ncols, 0, Wn.zero_rec, next0, win_pass_blocks(W0, next0, nb0));
        win_emit(P, ring, mask, M, Wn.zero_rec, img, W0, next0, nb0, r, s0, valid, true);
    }
    if (P.n_obs) {
        WinPart W1{P.obs, P.obs_stride, P.n_obs, P.num_det, 0, 0};
        const uint32_t nb1 = (P.n_obs + 31) / 32;
        win_records(P, r, W1.ncols, W1.col0, Wn.zero_rec, 0, win_pass_blocks(W1, 0, nb1));
        win_emit(P, ring, mask, M, Wn.zero_rec, img, W1, 0, nb1, r, s0, valid, true);
    }
}

}  // namespace

namespace sa {

size_t m2d_win_lds_bytes(uint32_t ring_words, uint32_t nchunks) {
    return (size_t)ring_words * 4 + 32 * WIN_PITCH + ((size_t)nchunks + 3) / 4 * 16;
}

hipError_t launch_m2d_win(const M2dParams &P, const M2dWindow &Wn, hipStream_t stream) {
    const uint64_t tiles = (P.shots + 31) / 32;
    if (tiles == 0) return hipSuccess;
    const size_t lds = m2d_win_lds_bytes(Wn.ring_words, Wn.nchunks);
    static size_t reserved = 0;
    const hipError_t e = reserve_lds(reinterpret_cast<const void *>(&m2d_win_kernel), lds, reserved);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(m2d_win_kernel, dim3((uint32_t)tiles), dim3(NTW), lds, stream, P, Wn);
    return hipGetLastError();
}

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
