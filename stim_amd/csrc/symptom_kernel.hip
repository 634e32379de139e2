// Device side of the sparse-symptom engine (see symptom.h).
//
// sym_sample_kernel: one workgroup per tile of S shots. The tile's noise positions are
// (site, shot) pairs over all site classes, concatenated class after class; lane t walks the
// contiguous range [t*R, (t+1)*R) drawing geometric gaps with the class's probability (restarting
// at class boundaries, exact because gaps are memoryless), picks each hit's outcome from the class's
// disjoint distribution, and XORs the outcome's basis symptoms into an LDS image of the tile's
// shot-major output rows (byte-exact copy of the destination rows, bit m of shot s at byte m/8, bit
// m%8 as in src/stim/py/numpy.pybind.cc:20-63). The image is then streamed to HBM with 16-byte stores.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "program.h"
#include "bitops.h"
#include "rng.h"
#include "launch_util.h"
#include "symptom.h"

using namespace sa;

namespace {

constexpr int NT = 256;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t TAG_SYM = 5u << 28;
constexpr uint32_t TAG_SCAN = 6u << 28;

__device__ __forceinline__ uint64_t gap(uint32_t r, float inv_log2_1mp) {
    const float u = ((float)(r >> 8) + 1.0f) * (1.0f / 16777216.0f);
    const float g = __log2f(u) * inv_log2_1mp;
    if (!(g < 1.0e12f)) return (uint64_t)1 << 40;
    return (uint64_t)g;
}

__device__ __forceinline__ uint32_t outcome(const uint32_t *otab, uint32_t woff, uint32_t nout_flags, uint32_t r) {
    const uint32_t nout = nout_flags & 0xFFFFu;
    if (nout_flags & SYM_UNIFORM) return otab[woff + 2 * __umulhi(r, nout) + 1];
    for (uint32_t i = 0; i + 1 < nout; i++) {
        if (r <= otab[woff + 2 * i]) return otab[woff + 2 * i + 1];
    }
    return otab[woff + 2 * (nout - 1) + 1];
}

struct Images {
    uint32_t *det;   // dets image (u32 words)
    uint32_t *obs;   // separate observable image
};

__device__ __forceinline__ void xor_entry(const SymDevParams &P, const Images &im, uint32_t m, uint32_t shot) {
    uint32_t bit;
    uint32_t *img;
    if (m < P.D) {
        bit = shot * (P.det_row_bytes * 8) + P.dets.col0 + m;
        img = im.det;
    } else if (P.obs_mode == 2) {
        bit = shot * (P.det_row_bytes * 8) + P.obs_col0 + (m - P.D);
        img = im.det;
    } else if (P.obs_mode == 1) {
        bit = shot * (P.obs_row_bytes * 8) + (m - P.D);
        img = im.obs;
    } else {
        return;
    }
    atomicXor(&img[bit >> 5], 1u << (bit & 31));
}

// XORs one column record's symptom into shot `shot` of the images.
__device__ __forceinline__ void xor_record(const SymDevParams &P, const Images &im, uint2 rec, uint32_t shot) {
    const uint32_t c = rec.x >> 28;
    if (c == 3) {
        const uint32_t off = rec.x & 0x0FFFFFFFu;
        for (uint32_t k = 0; k < rec.y; k++) xor_entry(P, im, P.ovf[off + k], shot);
        return;
    }
    if (c >= 1) xor_entry(P, im, rec.x & 0x0FFFFFFFu, shot);
    if (c == 2) xor_entry(P, im, rec.y, shot);
}

__device__ __forceinline__ void xor_column(const SymDevParams &P, const Images &im, uint32_t col, uint32_t shot) {
    xor_record(P, im, P.recs[col], shot);
}

// Applies the outcomes of two hits: loads every basis record of both hits first (independent
// loads in flight together), then XORs them into the images. Classes with more than 4 basis
// columns (correlated chains) take the one-at-a-time path.
__device__ __forceinline__ void apply_two(const SymDevParams &P, const Images &im, const SymClass &cl, uint32_t code1,
                                          uint32_t col1, uint32_t shot1, uint32_t code2, uint32_t col2,
                                          uint32_t shot2) {
    if (cl.basis > 4) {
        while (code1) {
            const uint32_t j = __ffs(code1) - 1;
            code1 &= code1 - 1;
            xor_column(P, im, col1 + j, shot1);
        }
        while (code2) {
            const uint32_t j = __ffs(code2) - 1;
            code2 &= code2 - 1;
            xor_column(P, im, col2 + j, shot2);
        }
        return;
    }
    uint2 r1[4], r2[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        r1[j] = (code1 >> j) & 1 ? P.recs[col1 + j] : make_uint2(0, 0);
        r2[j] = (code2 >> j) & 1 ? P.recs[col2 + j] : make_uint2(0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if ((code1 >> j) & 1) xor_record(P, im, r1[j], shot1);
        if ((code2 >> j) & 1) xor_record(P, im, r2[j], shot2);
    }
}

// Copies `rows` image rows (pitch `pitch` bytes) to dst rows of stride `stride`.
__device__ void write_rows(const uint8_t *img, uint32_t pitch, uint32_t rows, uint8_t *dst, uint64_t stride) {
    const uint32_t tid = threadIdx.x;
    if (stride == pitch) {
        const uint32_t total = rows * pitch;
        const uintptr_t a = (uintptr_t)dst;
        if ((a & 15) == 0) {
            const uint32_t n16 = total / 16;
            for (uint32_t i = tid; i < n16; i += NT) {
                const u32x4 v = *reinterpret_cast<const u32x4 *>(img + 16 * i);
                __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(dst + 16 * i));
            }
            for (uint32_t i = 16 * n16 + tid; i < total; i += NT) dst[i] = img[i];
        } else {
            for (uint32_t i = tid; i < total; i += NT) dst[i] = img[i];
        }
    } else {
        for (uint32_t i = tid; i < rows * pitch; i += NT) {
            const uint32_t r = i / pitch, c = i % pitch;
            dst[r * stride + c] = img[i];
        }
    }
}

// Hits queued per lane per round before the workgroup applies them together.
constexpr uint32_t HITS_PER_ROUND = 8;

__global__ __launch_bounds__(NT) void sym_sample_kernel(SymDevParams P) {
    extern __shared__ __align__(16) uint32_t lds[];
    const uint32_t S = P.shots_per_tile;
    const uint32_t det_words = (S * P.det_row_bytes + 15) / 16 * 4;
    const uint32_t obs_words = (S * P.obs_row_bytes + 15) / 16 * 4;
    Images im{lds, lds + det_words};
    uint2 *const queue = reinterpret_cast<uint2 *>(lds + det_words + obs_words);  // NT*HITS_PER_ROUND
    uint32_t *const qcount = lds + det_words + obs_words + 2 * NT * HITS_PER_ROUND;
    uint32_t *const otab_lds = qcount + 4;
    const bool otab_in_lds = P.otab_words <= SYM_OTAB_LDS_WORDS;
    const uint32_t *otab = otab_in_lds ? otab_lds : P.otab;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < det_words + obs_words; i += NT) lds[i] = 0;
    if (otab_in_lds) {
        for (uint32_t i = tid; i < P.otab_words; i += NT) otab_lds[i] = P.otab[i];
    }
    if (tid == 0) qcount[0] = 0;

    const uint64_t s0 = (uint64_t)blockIdx.x * S;
    const uint64_t tile64 = P.tile0 + blockIdx.x;
    const uint32_t tile = (uint32_t)tile64, lane_word = sym_lane_word(tid, tile64);
    const uint32_t k0 = (uint32_t)P.seed, k1 = (uint32_t)(P.seed >> 32);
    const uint64_t total = P.total_sites << P.tile_log2;
    const uint64_t R = (total + NT - 1) / NT;
    uint64_t pos = (uint64_t)tid * R;
    const uint64_t hi = pos + R < total ? pos + R : total;
    // This lane's walk state: class c with bounds [cbeg, cend) and the Philox counter.
    uint32_t c = 0, ctr = 0;
    SymClass cl{};
    uint64_t cbeg = 0, cend = 0;
    bool active = pos < hi;
    if (active) {
        uint32_t lo_c = 0, hi_c = P.nclasses;
        const uint64_t site_at = pos >> P.tile_log2;
        while (hi_c - lo_c > 1) {
            const uint32_t mid = (lo_c + hi_c) / 2;
            if (P.classes[mid].site0 <= site_at) lo_c = mid;
            else hi_c = mid;
        }
        c = lo_c;
        cl = P.classes[c];
        cbeg = (uint64_t)cl.site0 << P.tile_log2;
        const uint64_t ce = cbeg + ((uint64_t)cl.nsites << P.tile_log2);
        cend = ce < hi ? ce : hi;
    }
    __syncthreads();

    // Rounds: every lane queues up to HITS_PER_ROUND hits (walking geometric gaps, two per Philox
    // block), then all lanes apply the queued hits evenly.
    while (__syncthreads_or(active)) {
        uint32_t queued = 0;
        while (active && queued + 2 <= HITS_PER_ROUND) {
            if (pos >= cend || !(cl.p > 0.0f)) {
                // Next class (restart the gap walk at its first position).
                pos = cend;
                c++;
                if (pos >= hi || c >= P.nclasses) {
                    active = false;
                    break;
                }
                cl = P.classes[c];
                cbeg = (uint64_t)cl.site0 << P.tile_log2;
                const uint64_t ce = cbeg + ((uint64_t)cl.nsites << P.tile_log2);
                cend = ce < hi ? ce : hi;
                continue;
            }
            const uint4 r = philox10(make_uint4(ctr++, lane_word, tile, TAG_SYM), k0, k1);
            const uint64_t p1 = pos + gap(r.x, cl.inv_log);
            if (p1 >= cend) {
                pos = cend;
                continue;
            }
            const uint64_t p2 = p1 + 1 + gap(r.z, cl.inv_log);
            uint32_t slot = atomicAdd(qcount, p2 < cend ? 2u : 1u);
            {
                const uint64_t l = p1 - cbeg;
                queue[slot] = make_uint2(cl.col0 + (uint32_t)(l >> P.tile_log2) * cl.basis,
                                         (uint32_t)(l & (S - 1)) | (outcome(otab, cl.woff, cl.nout, r.y) << 10));
            }
            queued++;
            if (p2 < cend) {
                const uint64_t l = p2 - cbeg;
                queue[slot + 1] = make_uint2(cl.col0 + (uint32_t)(l >> P.tile_log2) * cl.basis,
                                             (uint32_t)(l & (S - 1)) | (outcome(otab, cl.woff, cl.nout, r.w) << 10));
                queued++;
                pos = p2 + 1;
            } else {
                pos = cend;
            }
        }
        __syncthreads();
        const uint32_t n = qcount[0];
        for (uint32_t i = tid; i < n; i += NT) {
            const uint2 h = queue[i];
            const uint32_t shot = h.y & 1023u;
            uint32_t code = h.y >> 10;
            if (code < 16) {
                uint2 rec[4];
#pragma unroll
                for (int j = 0; j < 4; j++) rec[j] = (code >> j) & 1 ? P.recs[h.x + j] : make_uint2(0, 0);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if ((code >> j) & 1) xor_record(P, im, rec[j], shot);
                }
            } else {
                while (code) {
                    const uint32_t j = __ffs(code) - 1;
                    code &= code - 1;
                    xor_column(P, im, h.x + j, shot);
                }
            }
        }
        __syncthreads();
        if (tid == 0) qcount[0] = 0;
    }
    __syncthreads();

    const uint64_t valid = P.shots > s0 ? (P.shots - s0 < S ? P.shots - s0 : S) : 0;
    if (P.dets.enabled)
        write_rows(reinterpret_cast<const uint8_t *>(im.det), P.det_row_bytes, (uint32_t)valid,
                   P.dets.ptr + s0 * P.dets.stride, P.dets.stride);
    if (P.obs.enabled && P.obs_row_bytes)
        write_rows(reinterpret_cast<const uint8_t *>(im.obs), P.obs_row_bytes, (uint32_t)valid,
                   P.obs.ptr + s0 * P.obs.stride, P.obs.stride);
}

// Scan formulation of the sparse draw (the workhorse sampler). For each sampling group (classes of
// equal probability, adjacent in site order), the tile's N = sites * S positions carry a Bernoulli(p)
// process; its hits are the partial sums of iid geometric gaps (the RareErrorIterator recurrence,
// probability_util.h:33-71). Every lane draws the same number m of gaps per round (no divergence), a
// workgroup prefix scan over the lanes' gap sums lays the concatenated gap sequence along the
// positions, hits that land inside [0, N) are applied, and rounds repeat until the sequence passes N.
// m is chosen so one round covers N with high probability. Streams are keyed by (tile, group, round,
// lane).
constexpr uint32_t SCAN_M_MAX = 8;

// Gaps are clamped at 2^31, above any group's position count N (at most 2^31 - 1), so a clamped gap
// always passes the end of the group and g + 1 cannot overflow.
__device__ __forceinline__ uint32_t gap32(uint32_t r, float inv_log2_1mp) {
    const float u = ((float)(r >> 8) + 1.0f) * (1.0f / 16777216.0f);
    const float g = __log2f(u) * inv_log2_1mp;
    return g < 2147483648.0f ? (uint32_t)g : 0x80000000u;
}

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    const uint32_t c = a + b;
    return c < a ? 0xFFFFFFFFu : c;
}

// Saturating prefix sum over the 256 lanes of the workgroup (saturating addition of non-negative
// values is associative, so any sum that reaches 2^32 - 1 stays there). Returns (exclusive, total).
__device__ __forceinline__ uint2 block_scan(uint32_t v, uint32_t *scratch) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x = sat_add(x, y);
    }
    if (lane == 63) scratch[wave] = x;
    uint32_t ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = 0;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < NT / 64; w++) {
        const uint32_t t = scratch[w];
        if (w < wave) before = sat_add(before, t);
        total = sat_add(total, t);
    }
    __syncthreads();
    return make_uint2(sat_add(before, ex), total);
}

// Where a hit's symptom entries land in a tile's LDS image. Entry m < D (a detector) of shot s is bit
// s * det_bits + det_col0 + m; entry m >= D (observable m - D) is bit obs_base + s * obs_bits +
// obs_col0 + (m - D), or is dropped when observables are not written. Bits are numbered across the
// workgroup's LDS (det_col0 and obs_base include the image's LDS address times 8, a multiple of 32),
// so bit b is bit b % 32 of the word at LDS byte address b / 32 * 4, reached without adding the image's
// base. The per-shot parts are computed once per hit (dbase, obase), so an entry costs a select, a
// shift and one LDS XOR, with no branch.
typedef __attribute__((address_space(3))) uint32_t lds_u32;

__device__ __forceinline__ uint32_t lds_address(uint32_t *p) {
    return (uint32_t)(uintptr_t)(lds_u32 *)p;
}

// XORs bit b (numbered across LDS) into its word.
__device__ __forceinline__ void lds_xor_bit(uint32_t b, uint32_t v) {
    __hip_atomic_fetch_xor((lds_u32 *)(uintptr_t)((b >> 3) & ~3u), v << (b & 31), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct Sink {
    uint32_t D;
    uint32_t det_bits, det_col0;
    uint32_t obs_base, obs_bits, obs_col0;
    bool obs_on;
};

__device__ __forceinline__ Sink make_sink(const SymDevParams &P, uint32_t *lds, uint32_t det_words) {
    Sink k;
    const uint32_t base = lds_address(lds) * 8;
    k.D = P.D;
    k.det_bits = P.det_row_bytes * 8;
    k.det_col0 = base + P.dets.col0;
    k.obs_on = P.obs_mode != 0;
    if (P.obs_mode == 2) {
        k.obs_base = base;
        k.obs_bits = k.det_bits;
        k.obs_col0 = P.obs_col0;
    } else {
        k.obs_base = base + det_words * 32;
        k.obs_bits = P.obs_row_bytes * 8;
        k.obs_col0 = 0;
    }
    return k;
}

__device__ __forceinline__ void sink_bit(const Sink &k, uint32_t m, uint32_t dbase, uint32_t obase) {
    const bool det = m < k.D;
    const uint32_t bit = (det ? dbase : obase) + m;
    if (det || k.obs_on) lds_xor_bit(bit, 1u);
}

__device__ __forceinline__ void sink_record(const Sink &k, const uint32_t *ovf, uint2 rec, uint32_t dbase,
                                            uint32_t obase) {
    const uint32_t c = rec.x >> 28;
    if (c == 3) {
        const uint32_t off = rec.x & 0x0FFFFFFFu;
        for (uint32_t e = 0; e < rec.y; e++) sink_bit(k, ovf[off + e], dbase, obase);
        return;
    }
    if (c >= 1) sink_bit(k, rec.x & 0x0FFFFFFFu, dbase, obase);
    if (c == 2) sink_bit(k, rec.y, dbase, obase);
}

__device__ __forceinline__ void sink_hit(const SymDevParams &P, const Sink &k, uint32_t colbase, uint32_t code,
                                         uint32_t shot, uint32_t basis) {
    const uint32_t dbase = shot * k.det_bits + k.det_col0;
    const uint32_t obase = k.obs_base + shot * k.obs_bits + k.obs_col0 - k.D;
    if (basis <= 4) {
        uint2 rec[4];
#pragma unroll
        for (int j = 0; j < 4; j++) rec[j] = (code >> j) & 1 ? P.recs[colbase + j] : make_uint2(0, 0);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if ((code >> j) & 1) sink_record(k, P.ovf, rec[j], dbase, obase);
        }
    } else {
        while (code) {
            const uint32_t j = __ffs(code) - 1;
            code &= code - 1;
            sink_record(k, P.ovf, P.recs[colbase + j], dbase, obase);
        }
    }
}

// Compact records (4 bytes; see symptom.h): detector entries go straight to dbase + m; the observable
// flags to observable bits obase + D + k (skipped when observables are not written).
__device__ __forceinline__ void xor_bit(const Sink &k, uint32_t bit) {
    lds_xor_bit(bit, 1u);
}

__device__ __forceinline__ void sink_record32(const SymDevParams &P, const Sink &k, uint32_t rec, uint32_t dbase,
                                              uint32_t obase) {
    const uint32_t c = (rec >> 28) & 3u;
    if (P.ovf32_any && c == 3) {
        const uint32_t off = rec & 0x0FFFFFFFu;
        const uint32_t n = P.ovf32[off];
        for (uint32_t e = 0; e < n; e++) sink_bit(k, P.ovf32[off + 1 + e], dbase, obase);
        return;
    }
    if (c >= 1) xor_bit(k, dbase + (rec & 0x3FFFu));
    if (c == 2) xor_bit(k, dbase + ((rec >> 14) & 0x3FFFu));
    if (k.obs_on && (rec >> 30)) {
        if (rec & (1u << 30)) xor_bit(k, obase + k.D);
        if (rec & (1u << 31)) xor_bit(k, obase + k.D + 1);
    }
}

// Applies one hit with compact records: the site's basis records start at `slot`, `rstride` apart
// sites; a single 16-byte load fetches all of them when rstride is 4 (classes with more than 4 basis
// columns load the records they use one by one).
// thin: the group is thinned (the site's threshold is slot 1; the hit is dropped when r exceeds it).
__device__ __forceinline__ bool sink_slot32(const SymDevParams &P, const Sink &k, uint32_t slot, uint32_t rstride,
                                            uint32_t code, uint32_t shot, bool thin, uint32_t r) {
    const uint32_t dbase = shot * k.det_bits + k.det_col0;
    const uint32_t obase = k.obs_base + shot * k.obs_bits + k.obs_col0 - k.D;
    if (rstride == 4) {
        const uint4 v = *reinterpret_cast<const uint4 *>(P.recs32 + slot);
        if (thin && r > v.y) return false;
        if (code & 1) sink_record32(P, k, v.x, dbase, obase);
        if (code & 2) sink_record32(P, k, v.y, dbase, obase);
        if (code & 4) sink_record32(P, k, v.z, dbase, obase);
        if (code & 8) sink_record32(P, k, v.w, dbase, obase);
    } else {
        while (code) {
            const uint32_t j = __ffs(code) - 1;
            code &= code - 1;
            sink_record32(P, k, P.recs32[slot + j], dbase, obase);
        }
    }
    return true;
}

// Marks mechanism `mech` as fired in row `row` of the launch's error rows.
__device__ __forceinline__ void record_error(const SymDevParams &P, uint64_t row, uint32_t mech) {
    atomicOr(&P.err_words[row * P.err_stride_words + (mech >> 5)], 1u << (mech & 31));
}

// Saturating prefix sum over the 64 lanes of a wave. Returns (exclusive, total).
__device__ __forceinline__ uint2 wave_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x = sat_add(x, y);
    }
    const uint32_t total = __shfl(x, 63, 64);
    uint32_t ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = 0;
    return make_uint2(ex, total);
}

// Position where draw-space coordinate a falls in a group that starts at draw-space coordinate c and
// has N positions at probability p (draw space: expected draws, accumulated over the groups in
// order). Adjacent waves evaluate the same expression for their shared edge, so their ranges meet.
__device__ __forceinline__ uint32_t draw_position(float a, float c, float p, uint32_t N) {
    if (a <= c) return 0;
    const float x = (a - c) / p;
    return x >= (float)N ? N : (uint32_t)x;
}

// One tile of the scan sampler with the groups' positions shared among the waves (DEM samplers, whose
// many probability buckets would each cost a workgroup scan round): the groups' expected draws are
// laid end to end and cut into NT / 64 equal shares, one per wave; a wave draws every group's range
// that falls in its share, restarting the process at each range's start (exact, since geometric gaps
// are memoryless), with prefix scans within the wave and no workgroup barrier between the image's
// zeroing and its write-out. Streams are keyed by (tile, group, round, lane), as in scan_tile.
__device__ __forceinline__ void scan_tile_waves(const SymDevParams &P, uint32_t btile, uint32_t *lds) {
    const uint32_t S = P.shots_per_tile;
    const uint32_t det_words = (S * P.det_row_bytes + 15) / 16 * 4;
    const uint32_t obs_words = (S * P.obs_row_bytes + 15) / 16 * 4;
    uint32_t *const otab_lds = lds + det_words + obs_words + 8;
    const bool otab_in_lds = P.otab_words <= SYM_OTAB_LDS_WORDS;
    const uint32_t *otab = otab_in_lds ? otab_lds : P.otab;
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    {
        uint4 *const v = reinterpret_cast<uint4 *>(lds);
        for (uint32_t i = tid; i < (det_words + obs_words) / 4; i += NT) v[i] = make_uint4(0, 0, 0, 0);
    }
    if (otab_in_lds) {
        for (uint32_t i = tid; i < P.otab_words; i += NT) otab_lds[i] = P.otab[i];
    }
    const Sink sink = make_sink(P, lds, det_words);
    float total = 0.0f;
    for (uint32_t gi = 0; gi < P.ngroups; gi++) {
        const SymGroup G = P.groups[gi];
        if (G.p > 0.0f) total += (float)(G.nsites << P.tile_log2) * G.p;
    }
    const float a0 = total * (float)wave / (float)(NT / 64), a1 = total * (float)(wave + 1) / (float)(NT / 64);
    __syncthreads();

    const uint64_t s0 = (uint64_t)btile * S;
    const uint64_t tile64 = P.tile0 + btile;
    const uint32_t tile = (uint32_t)tile64, lane_word = sym_lane_word(tid, tile64);
    const uint32_t k0 = (uint32_t)P.seed, k1 = (uint32_t)(P.seed >> 32);
    float c = 0.0f;
    for (uint32_t gi = 0; gi < P.ngroups; gi++) {
        const SymGroup G = P.groups[gi];
        if (!(G.p > 0.0f)) continue;
        const uint32_t Ng = G.nsites << P.tile_log2;
        const float h = (float)Ng * G.p;
        const uint32_t lo = draw_position(a0, c, G.p, Ng);
        const uint32_t hi = wave + 1 == NT / 64 ? Ng : draw_position(a1, c, G.p, Ng);
        c += h;
        if (hi <= lo) continue;
        const uint32_t N = hi - lo;
        const float mu = (float)N * G.p;
        uint32_t m = (uint32_t)((mu + 3.0f * sqrtf(mu) + 4.0f) / 64) + 1;
        if (m > SCAN_M_MAX) m = SCAN_M_MAX;
        uint32_t base = 0;
        for (uint32_t round = 0; base < N; round++) {
            uint32_t g[SCAN_M_MAX], oc[SCAN_M_MAX];
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX / 2; k++) {
                if (2 * k < m) {
                    const uint4 r = philox10(make_uint4(round * (SCAN_M_MAX / 2) + k, lane_word, tile, TAG_SCAN | gi), k0, k1);
                    g[2 * k] = gap32(r.x, G.inv_log);
                    oc[2 * k] = r.y;
                    g[2 * k + 1] = gap32(r.z, G.inv_log);
                    oc[2 * k + 1] = r.w;
                } else {
                    g[2 * k] = g[2 * k + 1] = 0;
                    oc[2 * k] = oc[2 * k + 1] = 0;
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX; k++) {
                if (k < m) sum = sat_add(sum, g[k] + 1);
            }
            const uint2 sc = wave_scan(sum);
            uint32_t pos = sat_add(base, sc.x);
            pos = pos < N ? pos : N;
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX; k++) {
                if (k < m) {
                    pos += g[k] < N ? g[k] : N;
                    pos = pos < N ? pos : N;
                    if (pos < N) {
                        const uint32_t srel = (lo + pos) >> P.tile_log2;
                        const uint32_t shot = (lo + pos) & (S - 1);
                        // Member class of the site (a group holds up to 4 classes in site order).
                        const uint32_t mk = (srel >= G.bound[1] ? 1u : 0u) + (srel >= G.bound[2] ? 1u : 0u) +
                                            (srel >= G.bound[3] ? 1u : 0u);
                        const uint32_t nf = mk == 0 ? G.nout[0] : mk == 1 ? G.nout[1] : mk == 2 ? G.nout[2] : G.nout[3];
                        uint32_t code;
                        if (nf & SYM_SEQ) {
                            code = __umulhi(oc[k], nf & 0xFFFFu) + 1;
                        } else {
                            const uint32_t wo =
                                mk == 0 ? G.woff[0] : mk == 1 ? G.woff[1] : mk == 2 ? G.woff[2] : G.woff[3];
                            code = outcome(otab, wo, nf, oc[k]);
                        }
                        if (G.flags & SYM_DEM_SLOTS) code = SYM_DEM_SLOT_CODE;
                        const bool thin = (G.flags & SYM_THIN) != 0;
                        bool fired;
                        if (P.recs32) {
                            fired = sink_slot32(P, sink, G.rec0 + srel * G.rstride, G.rstride, code, shot, thin, oc[k]);
                        } else {
                            fired = !thin || oc[k] <= P.thr[G.site0 + srel];
                            if (fired) {
                                const SymClass cl = P.classes[G.c0];
                                sink_hit(P, sink, cl.col0 + srel * cl.basis, code, shot, cl.basis);
                            }
                        }
                        if (P.err_words && fired && s0 + shot < P.shots)
                            record_error(P, s0 + shot, P.site_mech[G.site0 + srel]);
                    }
                    pos += 1;
                }
            }
            base = sat_add(base, sc.y);
            base = base < N ? base : N;
        }
    }
    __syncthreads();

    const uint64_t valid = P.shots > s0 ? (P.shots - s0 < S ? P.shots - s0 : S) : 0;
    if (P.dets.enabled)
        write_rows(reinterpret_cast<const uint8_t *>(lds), P.det_row_bytes, (uint32_t)valid,
                   P.dets.ptr + s0 * P.dets.stride, P.dets.stride);
    if (P.obs.enabled && P.obs_row_bytes)
        write_rows(reinterpret_cast<const uint8_t *>(lds + det_words), P.obs_row_bytes, (uint32_t)valid,
                   P.obs.ptr + s0 * P.obs.stride, P.obs.stride);
}

// One tile of the scan sampler: tile `btile` of the launch described by P, in the workgroup's LDS.
__device__ __forceinline__ void scan_tile(const SymDevParams &P, uint32_t btile, uint32_t *lds) {
    const uint32_t S = P.shots_per_tile;
    const uint32_t det_words = (S * P.det_row_bytes + 15) / 16 * 4;
    const uint32_t obs_words = (S * P.obs_row_bytes + 15) / 16 * 4;
    Images im{lds, lds + det_words};
    uint32_t *const scratch = lds + det_words + obs_words;  // 8 words
    uint32_t *const otab_lds = scratch + 8;
    const bool otab_in_lds = P.otab_words <= SYM_OTAB_LDS_WORDS;
    const uint32_t *otab = otab_in_lds ? otab_lds : P.otab;
    const uint32_t tid = threadIdx.x;
    {
        // The image is a whole number of 16-byte words.
        uint4 *const v = reinterpret_cast<uint4 *>(lds);
        for (uint32_t i = tid; i < (det_words + obs_words) / 4; i += NT) v[i] = make_uint4(0, 0, 0, 0);
    }
    if (otab_in_lds) {
        for (uint32_t i = tid; i < P.otab_words; i += NT) otab_lds[i] = P.otab[i];
    }
    const Sink sink = make_sink(P, lds, det_words);
    __syncthreads();

    const uint64_t s0 = (uint64_t)btile * S;
    const uint64_t tile64 = P.tile0 + btile;
    const uint32_t tile = (uint32_t)tile64, lane_word = sym_lane_word(tid, tile64);
    const uint32_t k0 = (uint32_t)P.seed, k1 = (uint32_t)(P.seed >> 32);
    for (uint32_t gi = 0; gi < P.ngroups; gi++) {
        const SymGroup G = P.groups[gi];
        if (!(G.p > 0.0f)) continue;
        const uint32_t N = G.nsites << P.tile_log2;
        // Gaps per lane per round: cover the expected hits plus three standard deviations (a short
        // extra round follows in the rare tiles that need it).
        const float mu = (float)N * G.p;
        uint32_t m = (uint32_t)((mu + 3.0f * sqrtf(mu) + 4.0f) / NT) + 1;
        if (m > SCAN_M_MAX) m = SCAN_M_MAX;
        uint32_t base = 0;
        for (uint32_t round = 0; base < N; round++) {
            uint32_t g[SCAN_M_MAX], oc[SCAN_M_MAX];
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX / 2; k++) {
                if (2 * k < m) {
                    const uint4 r = philox10(make_uint4(round * (SCAN_M_MAX / 2) + k, lane_word, tile, TAG_SCAN | gi), k0, k1);
                    g[2 * k] = gap32(r.x, G.inv_log);
                    oc[2 * k] = r.y;
                    g[2 * k + 1] = gap32(r.z, G.inv_log);
                    oc[2 * k + 1] = r.w;
                } else {
                    g[2 * k] = g[2 * k + 1] = 0;
                    oc[2 * k] = oc[2 * k + 1] = 0;
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX; k++) {
                if (k < m) sum = sat_add(sum, g[k] + 1);
            }
            const uint2 sc = block_scan(sum, scratch);
            // Positions stay at most N < 2^31: the start and every gap are clamped at N.
            uint32_t pos = sat_add(base, sc.x);
            pos = pos < N ? pos : N;
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX; k++) {
                if (k < m) {
                    pos += g[k] < N ? g[k] : N;
                    pos = pos < N ? pos : N;
                    if (pos < N) {
                        const uint32_t srel = pos >> P.tile_log2;
                        const uint32_t shot = pos & (S - 1);
                        // Member class of the site (a group holds up to 4 classes in site order).
                        const uint32_t mk = (srel >= G.bound[1] ? 1u : 0u) + (srel >= G.bound[2] ? 1u : 0u) +
                                            (srel >= G.bound[3] ? 1u : 0u);
                        const uint32_t nf = mk == 0 ? G.nout[0] : mk == 1 ? G.nout[1] : mk == 2 ? G.nout[2] : G.nout[3];
                        uint32_t code;
                        if (nf & SYM_SEQ) {
                            code = __umulhi(oc[k], nf & 0xFFFFu) + 1;
                        } else {
                            const uint32_t wo =
                                mk == 0 ? G.woff[0] : mk == 1 ? G.woff[1] : mk == 2 ? G.woff[2] : G.woff[3];
                            code = outcome(otab, wo, nf, oc[k]);
                        }
                        if (G.flags & SYM_DEM_SLOTS) code = SYM_DEM_SLOT_CODE;
                        const bool thin = (G.flags & SYM_THIN) != 0;
                        bool fired;
                        if (P.recs32) {
                            fired = sink_slot32(P, sink, G.rec0 + srel * G.rstride, G.rstride, code, shot, thin, oc[k]);
                        } else {
                            fired = !thin || oc[k] <= P.thr[G.site0 + srel];
                            if (fired) {
                                const SymClass cl = P.classes[G.c0];
                                sink_hit(P, sink, cl.col0 + srel * cl.basis, code, shot, cl.basis);
                            }
                        }
                        if (P.err_words && fired && s0 + shot < P.shots)
                            record_error(P, s0 + shot, P.site_mech[G.site0 + srel]);
                    }
                    pos += 1;
                }
            }
            base = sat_add(base, sc.y);
            base = base < N ? base : N;
        }
    }
    __syncthreads();

    const uint64_t valid = P.shots > s0 ? (P.shots - s0 < S ? P.shots - s0 : S) : 0;
    if (P.dets.enabled)
        write_rows(reinterpret_cast<const uint8_t *>(im.det), P.det_row_bytes, (uint32_t)valid,
                   P.dets.ptr + s0 * P.dets.stride, P.dets.stride);
    if (P.obs.enabled && P.obs_row_bytes)
        write_rows(reinterpret_cast<const uint8_t *>(im.obs), P.obs_row_bytes, (uint32_t)valid,
                   P.obs.ptr + s0 * P.obs.stride, P.obs.stride);
}

// Applies up to K hits of one lane (the specialised scan kernel's pass 2 and 3): loads the K hits'
// site entries together, then their K outcome records together (entries past the round's hits name
// the group's first site and are skipped), so that the compiler can wait for each load in turn; then
// flips each hit's detector bits without branching per entry (an absent entry XORs 0 into its word)
// and its observable flags. site[k] = the hit's site entry index, oc[k] = its outcome draw: the code is
// mulhi(oc, nout) + 1 with the site's nout from its entry, as the generic kernel computes it.
template <bool OBS, uint32_t K>
__device__ __forceinline__ void apply_outcomes(const SymDevParams &P, const Sink &sink, const uint32_t *site,
                                               const uint32_t *oc, const uint32_t *shot, uint32_t S) {
    uint2 se[K], rec[K];
#pragma unroll
    for (uint32_t k = 0; k < K; k++) se[k] = P.orec_sites[site[k]];
#pragma unroll
    for (uint32_t k = 0; k < K; k++) rec[k] = P.orecs[(se[k].x & 0x0FFFFFFFu) + __umulhi(oc[k], se[k].x >> 28)];
#pragma unroll
    for (uint32_t k = 0; k < K; k++) {
        if (shot[k] < S) {
            // Shot and row-bit counts fit 24 bits (S <= 8192, rows below 64 KB); the site's base
            // detector is below D, so every entry's bit lies in the shot's row.
            const uint32_t dbase = __umul24(shot[k], sink.det_bits) + sink.det_col0 + se[k].y;
            const uint2 r = rec[k];
            // An overflow record holds its list offset where entries 2 and 3 would be (and no valid
            // bits): read those entries, and its observable flags, as 0 so that every address stays
            // inside the tile's image.
            const uint32_t ry = (r.y & SYM_OREC_OVERFLOW) ? 0u : r.y;
            const uint32_t f[4] = {r.x & 0x3FFFu, (r.x >> 14) & 0x3FFFu, ry & 0x3FFFu, (ry >> 14) & 0x3FFFu};
#pragma unroll
            for (uint32_t e = 0; e < 4; e++) lds_xor_bit(dbase + f[e], (r.x >> (28 + e)) & 1u);
            if (OBS && ((ry >> 28) & 3u)) {
                const uint32_t obase = sink.obs_base + __umul24(shot[k], sink.obs_bits) + sink.obs_col0;
                if ((ry >> 29) & 1u) xor_bit(sink, obase);
                if ((ry >> 28) & 1u) xor_bit(sink, obase + 1);
            }
            if (r.y & SYM_OREC_OVERFLOW) {
                const uint32_t obase = sink.obs_base + __umul24(shot[k], sink.obs_bits) + sink.obs_col0;
                // More than four detectors: the list (detector m, observable D + k).
                const uint32_t *list = P.orec_ovf + (r.y & ~SYM_OREC_OVERFLOW);
                const uint32_t cnt = list[0];
                for (uint32_t e = 0; e < cnt; e++) {
                    const uint32_t v = list[1 + e];
                    if (v < sink.D) xor_bit(sink, dbase + v);
                    else if (OBS) xor_bit(sink, obase + (v - sink.D));
                }
            }
        }
    }
}

// Specialised scan tile for models whose groups are plain (compact records, no thinning) and whose
// members draw outcomes uniformly over codes 1..nout: the QEC memory circuits. It draws exactly the
// hits scan_tile draws (same counters, gaps, positions and outcome codes), so its rows are
// bit-identical; it differs in how a hit is applied. Each (site, outcome) has one 8-byte outcome
// record holding the outcome's whole symptom (entries that cancel between the outcome's basis columns
// already removed), so a hit reads one record and flips at most four detector bits plus observable
// flags, instead of testing four basis records of two entries each. Each round first forms every
// hit's record index and shot, then issues all record loads together (a lane past the end of the
// group reads the group's first record and applies nothing), then applies the hits.
template <bool OBS>
__device__ __forceinline__ void scan_tile_fast(const SymDevParams &P, uint32_t btile, uint32_t *lds) {
    const uint32_t S = P.shots_per_tile;
    const uint32_t det_words = (S * P.det_row_bytes + 15) / 16 * 4;
    const uint32_t obs_words = (S * P.obs_row_bytes + 15) / 16 * 4;
    uint32_t *const scratch = lds + det_words + obs_words;  // 8 words
    const uint32_t tid = threadIdx.x;
    {
        uint4 *const v = reinterpret_cast<uint4 *>(lds);
        for (uint32_t i = tid; i < (det_words + obs_words) / 4; i += NT) v[i] = make_uint4(0, 0, 0, 0);
    }
    const Sink sink = make_sink(P, lds, det_words);
    __syncthreads();

    const uint64_t s0 = (uint64_t)btile * S;
    const uint64_t tile64 = P.tile0 + btile;
    const uint32_t tile = (uint32_t)tile64, lane_word = sym_lane_word(tid, tile64);
    const uint32_t k0 = (uint32_t)P.seed, k1 = (uint32_t)(P.seed >> 32);
    const uint32_t lg = P.tile_log2;
    for (uint32_t gi = 0; gi < P.ngroups; gi++) {
        const SymGroup G = P.groups[gi];
        if (!(G.p > 0.0f)) continue;
        const uint32_t N = G.nsites << lg;
        const float mu = (float)N * G.p;
        uint32_t m = (uint32_t)((mu + 3.0f * sqrtf(mu) + 4.0f) / NT) + 1;
        if (m > SCAN_M_MAX) m = SCAN_M_MAX;
        const uint32_t b0 = P.orec_base[4 * gi];
        uint32_t base = 0;
        for (uint32_t round = 0; base < N; round++) {
            uint32_t g[SCAN_M_MAX], oc[SCAN_M_MAX];
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX / 2; k++) {
                if (2 * k < m) {
                    const uint4 r = philox10(make_uint4(round * (SCAN_M_MAX / 2) + k, lane_word, tile, TAG_SCAN | gi), k0, k1);
                    g[2 * k] = gap32(r.x, G.inv_log);
                    oc[2 * k] = r.y;
                    g[2 * k + 1] = gap32(r.z, G.inv_log);
                    oc[2 * k + 1] = r.w;
                } else {
                    g[2 * k] = g[2 * k + 1] = 0;
                    oc[2 * k] = oc[2 * k + 1] = 0;
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX; k++) {
                if (k < m) sum = sat_add(sum, g[k] + 1);
            }
            const uint2 sc = block_scan(sum, scratch);
            uint32_t pos = sat_add(base, sc.x);
            pos = pos < N ? pos : N;
            // Pass 1: site entry and shot of every hit (shot S marks no hit; its entry is the group's
            // first site's, which is read and not applied).
            uint32_t idx[SCAN_M_MAX], shot[SCAN_M_MAX];
#pragma unroll
            for (uint32_t k = 0; k < SCAN_M_MAX; k++) {
                idx[k] = b0;
                shot[k] = S;
                if (k < m) {
                    pos += g[k] < N ? g[k] : N;
                    pos = pos < N ? pos : N;
                    if (pos < N) {
                        idx[k] = b0 + (pos >> lg);
                        shot[k] = pos & (S - 1);
                    }
                    pos += 1;
                }
            }
            // Pass 2 and 3: record loads in flight together, then the flips (m is uniform).
            static_assert(SCAN_M_MAX == 8, "two apply widths");
            if (m > 4) apply_outcomes<OBS, 8>(P, sink, idx, oc, shot, S);
            else apply_outcomes<OBS, 4>(P, sink, idx, oc, shot, S);
            base = sat_add(base, sc.y);
            base = base < N ? base : N;
        }
    }
    __syncthreads();

    const uint64_t valid = P.shots > s0 ? (P.shots - s0 < S ? P.shots - s0 : S) : 0;
    if (P.dets.enabled)
        write_rows(reinterpret_cast<const uint8_t *>(lds), P.det_row_bytes, (uint32_t)valid,
                   P.dets.ptr + s0 * P.dets.stride, P.dets.stride);
    if (P.obs.enabled && P.obs_row_bytes)
        write_rows(reinterpret_cast<const uint8_t *>(lds + det_words), P.obs_row_bytes, (uint32_t)valid,
                   P.obs.ptr + s0 * P.obs.stride, P.obs.stride);
}

// Both scan kernels ask for 6 waves per SIMD (at most 80 VGPRs, no spills): the sampler is bound by
// instruction issue and latency, so resident waves matter more than registers.
__global__ __launch_bounds__(NT, 6) void sym_scan_kernel(SymDevParams P) {
    extern __shared__ __align__(16) uint32_t lds[];
    scan_tile(P, blockIdx.x, lds);
}

// The wave-split form (P.wave_split: DEM samplers) in a kernel of its own.
__global__ __launch_bounds__(NT, 6) void sym_scan_waves_kernel(SymDevParams P) {
    extern __shared__ __align__(16) uint32_t lds[];
    scan_tile_waves(P, blockIdx.x, lds);
}

// The specialised form holds a round's record loads in registers (16 more VGPRs); at d=25 its
// 31 KB tile image allows 5 workgroups per CU, so it asks for 5 waves per SIMD.
template <bool OBS>
__global__ __launch_bounds__(NT, 5) void sym_scan_fast_kernel(SymDevParams P) {
    extern __shared__ __align__(16) uint32_t lds[];
    scan_tile_fast<OBS>(P, blockIdx.x, lds);
}

// A batch job's tile: the specialised scan for jobs whose sampler has outcome records (the job's
// plain launch runs that kernel too; rows are identical either way), the generic scan otherwise.
__device__ __forceinline__ void batch_tile(const SymDevParams &P, uint32_t btile, uint32_t *lds) {
    if (P.orecs) {
        if (P.obs_mode != 0) scan_tile_fast<true>(P, btile, lds);
        else scan_tile_fast<false>(P, btile, lds);
    } else if (P.wave_split) {
        scan_tile_waves(P, btile, lds);
    } else {
        scan_tile(P, btile, lds);
    }
}

// Batched form: many launches (jobs, possibly of different circuits) in one kernel; job j owns global
// tiles [prefix[j], prefix[j+1]). Persistent schedule (grid smaller than the tile count): workgroups
// claim chunks of BATCH_CHUNK consecutive tiles from a counter, caching the current job so the search
// only runs when a chunk crosses into another job, and leave once the counter passes the last tile.
// With a grid of one workgroup per tile, each workgroup samples its own tile.
constexpr uint32_t BATCH_CHUNK = 8;

__global__ __launch_bounds__(NT, 6) void sym_batch_kernel(const SymDevParams *__restrict__ jobs,
                                                       const uint64_t *__restrict__ prefix, uint32_t njobs,
                                                       uint32_t *counter) {
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ uint32_t next;
    const uint64_t total = prefix[njobs];
    if ((uint64_t)gridDim.x >= total) {
        // One tile per workgroup (the dispatcher schedules the tiles).
        const uint64_t t = blockIdx.x;
        uint32_t a = 0, b = njobs;
        while (b - a > 1) {
            const uint32_t mid = (a + b) / 2;
            if (prefix[mid] <= t) a = mid;
            else b = mid;
        }
        const uint32_t j = __builtin_amdgcn_readfirstlane(a);
        if (t < total) {
            const SymDevParams P = jobs[j];
            batch_tile(P, (uint32_t)(t - prefix[j]), lds);
        }
        return;
    }
    uint32_t j = 0;
    uint64_t lo = 0, hi = 0;   // tile range of job j (empty until the first search)
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) next = atomicAdd(counter, BATCH_CHUNK);
        __syncthreads();
        const uint64_t t0 = next;
        if (t0 >= total) break;
        const uint64_t t1 = t0 + BATCH_CHUNK < total ? t0 + BATCH_CHUNK : total;
        for (uint64_t t = t0; t < t1; t++) {
            if (t < lo || t >= hi) {
                uint32_t a = 0, b = njobs;
                while (b - a > 1) {
                    const uint32_t mid = (a + b) / 2;
                    if (prefix[mid] <= t) a = mid;
                    else b = mid;
                }
                j = __builtin_amdgcn_readfirstlane(a);
                lo = prefix[j];
                hi = prefix[j + 1];
            }
            if (t != t0) __syncthreads();   // the previous tile's rows are written before the image is reused
            const SymDevParams P = jobs[j];
            batch_tile(P, (uint32_t)(t - lo), lds);
        }
    }
}

// Batch bookkeeping kernels (kernels rather than memory operations, so the batch can be captured in a
// graph): reset the tile counter before a run; after it, advance every job's RNG tile base by the
// tiles its sampler used in the run (all of that sampler's jobs), as a plain call would.
__global__ void sym_batch_reset_kernel(uint32_t *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

__global__ void sym_batch_advance_kernel(SymDevParams *jobs, const uint64_t *advance, uint32_t njobs) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs) jobs[j].tile0 += advance[j];
}

// Fixed-fault mode of the symptom engine: shot s gets the XOR of the columns listed in
// cols[off[s] .. off[s+1]).
__global__ void sym_fixed_kernel(SymDevParams P, const uint64_t *__restrict__ off, const uint32_t *__restrict__ cols) {
    extern __shared__ __align__(16) uint32_t lds[];
    const uint32_t S = P.shots_per_tile;
    const uint32_t det_words = (S * P.det_row_bytes + 15) / 16 * 4;
    const uint32_t obs_words = (S * P.obs_row_bytes + 15) / 16 * 4;
    Images im{lds, lds + det_words};
    for (uint32_t i = threadIdx.x; i < det_words + obs_words; i += NT) lds[i] = 0;
    __syncthreads();
    const uint64_t s0 = (uint64_t)blockIdx.x * S;
    const uint64_t valid = P.shots > s0 ? (P.shots - s0 < S ? P.shots - s0 : S) : 0;
    for (uint32_t r = 0; r < valid; r++) {
        for (uint64_t i = off[s0 + r] + threadIdx.x; i < off[s0 + r + 1]; i += NT) xor_column(P, im, cols[i], r);
    }
    __syncthreads();
    if (P.dets.enabled)
        write_rows(reinterpret_cast<const uint8_t *>(im.det), P.det_row_bytes, (uint32_t)valid,
                   P.dets.ptr + s0 * P.dets.stride, P.dets.stride);
    if (P.obs.enabled && P.obs_row_bytes)
        write_rows(reinterpret_cast<const uint8_t *>(im.obs), P.obs_row_bytes, (uint32_t)valid,
                   P.obs.ptr + s0 * P.obs.stride, P.obs.stride);
}

// Counts set bits of each column's dets row + obs row.
// Symptom-table extraction from the fixed-fault rows (one row per basis fault), one wave per row: lanes
// read the row's bytes 64 at a time (coalesced); the count kernel sums their popcounts, the fill kernel
// places each lane's set bits after a wave prefix sum of the counts, so a row's entries come out in
// increasing bit order.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_exclusive(uint32_t v, uint32_t lane) {
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x - v;
}

__global__ __launch_bounds__(256) void sym_count_kernel(const uint8_t *__restrict__ det, uint32_t det_row,
                                                        const uint8_t *__restrict__ obs, uint32_t obs_row,
                                                        uint64_t rows, uint32_t *__restrict__ counts) {
    const uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= rows) return;   // whole waves leave together
    uint32_t c = 0;
    const uint8_t *d = det + r * det_row;
    for (uint32_t i = lane; i < det_row; i += 64) c += __popc(d[i]);
    const uint8_t *o = obs + r * obs_row;
    for (uint32_t i = lane; i < obs_row; i += 64) c += __popc(o[i]);
    c = wave_sum(c);
    if (lane == 0) counts[r] = c;
}

__global__ __launch_bounds__(256) void sym_fill_kernel(const uint8_t *__restrict__ det, uint32_t det_row,
                                                       const uint8_t *__restrict__ obs, uint32_t obs_row, uint64_t rows,
                                                       uint32_t D, const uint32_t *__restrict__ offs,
                                                       uint32_t *__restrict__ ent) {
    const uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= rows) return;
    uint32_t k = offs[r];
    for (int part = 0; part < 2; part++) {
        const uint8_t *src = part == 0 ? det + r * det_row : obs + r * obs_row;
        const uint32_t len = part == 0 ? det_row : obs_row;
        const uint32_t base_bit = part == 0 ? 0u : D;
        for (uint32_t c0 = 0; c0 < len; c0 += 64) {
            const uint32_t i = c0 + lane;
            uint32_t v = i < len ? src[i] : 0u;
            const uint32_t n = __popc(v);
            const uint32_t before = wave_exclusive(n, lane);
            const uint32_t total = __shfl(before + n, 63, 64);
            uint32_t at = k + before;
            while (v) {
                const uint32_t b = __ffs(v) - 1;
                v &= v - 1;
                ent[at++] = base_bit + 8 * i + b;
            }
            k += total;
        }
    }
}

// Packed rows -> one byte per entry (value bit + base_char).
__global__ void unpack_rows_kernel(const uint8_t *__restrict__ src, uint64_t src_stride, uint8_t *__restrict__ dst,
                                   uint64_t dst_stride, uint64_t rows, uint32_t ncols, uint32_t base_char) {
    const uint64_t total = rows * ncols;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = i / ncols;
        const uint32_t c = (uint32_t)(i % ncols);
        dst[r * dst_stride + c] = (uint8_t)(((src[r * src_stride + (c >> 3)] >> (c & 7)) & 1u) + base_char);
    }
}

// Philox4x32-10 of the samplers on n counters under one key (checks of the generator itself).
__global__ void philox_kernel(const uint4 *__restrict__ ctr, uint64_t n, uint32_t k0, uint32_t k1,
                              uint4 *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = philox10(ctr[i], k0, k1);
}

}  // namespace

namespace sa {

hipError_t launch_philox(const uint32_t *ctr, uint64_t n, uint64_t key, uint32_t *out, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(philox_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream,
                       reinterpret_cast<const uint4 *>(ctr), n, (uint32_t)key, (uint32_t)(key >> 32),
                       reinterpret_cast<uint4 *>(out));
    return hipGetLastError();
}

hipError_t launch_sym_sample(const SymDevParams &P, uint32_t tiles, size_t lds, hipStream_t stream) {
    if (!tiles) return hipSuccess;
    static size_t reserved = 0;
    const hipError_t e = reserve_lds((const void *)sym_sample_kernel, lds, reserved);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sym_sample_kernel, dim3(tiles), dim3(NT), lds, stream, P);
    return hipGetLastError();
}

hipError_t launch_sym_scan(const SymDevParams &P, uint32_t tiles, size_t lds, hipStream_t stream, bool fast) {
    if (!tiles) return hipSuccess;
    if (fast) {
        if (P.obs_mode != 0) {
            static size_t reserved_obs = 0;
            const hipError_t e = reserve_lds((const void *)sym_scan_fast_kernel<true>, lds, reserved_obs);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(sym_scan_fast_kernel<true>, dim3(tiles), dim3(NT), lds, stream, P);
        } else {
            static size_t reserved_plain = 0;
            const hipError_t e = reserve_lds((const void *)sym_scan_fast_kernel<false>, lds, reserved_plain);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(sym_scan_fast_kernel<false>, dim3(tiles), dim3(NT), lds, stream, P);
        }
        return hipGetLastError();
    }
    if (P.wave_split) {
        static size_t reserved_waves = 0;
        const hipError_t e = reserve_lds((const void *)sym_scan_waves_kernel, lds, reserved_waves);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(sym_scan_waves_kernel, dim3(tiles), dim3(NT), lds, stream, P);
        return hipGetLastError();
    }
    static size_t reserved = 0;
    const hipError_t e = reserve_lds((const void *)sym_scan_kernel, lds, reserved);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sym_scan_kernel, dim3(tiles), dim3(NT), lds, stream, P);
    return hipGetLastError();
}

// The dynamic LDS attribute is set by sym_batch_blocks_per_cu (before any graph capture).
// advance == null: no tile-base update (callers that set the bases themselves).
hipError_t launch_sym_batch(SymDevParams *jobs, const uint64_t *prefix, const uint64_t *advance, uint32_t njobs,
                            uint32_t *counter, uint32_t grid, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL(sym_batch_reset_kernel, dim3(1), dim3(64), 0, stream, counter);
    hipLaunchKernelGGL(sym_batch_kernel, dim3(grid), dim3(NT), lds, stream, jobs, prefix, njobs, counter);
    if (advance)
        hipLaunchKernelGGL(sym_batch_advance_kernel, dim3((njobs + 255) / 256), dim3(256), 0, stream, jobs, advance,
                           njobs);
    return hipGetLastError();
}

// Raises (never lowers) the batch kernel's dynamic-LDS limit to lds, then returns how many of its
// workgroups with that much LDS fit on one CU. Every launch size is announced through here first.
int sym_batch_blocks_per_cu(size_t lds) {
    static size_t reserved = 0;
    (void)reserve_lds((const void *)sym_batch_kernel, lds, reserved);
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sym_batch_kernel, NT, lds) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_sym_fixed(const SymDevParams &P, const uint64_t *off, const uint32_t *cols, uint32_t tiles,
                            size_t lds, hipStream_t stream) {
    if (!tiles) return hipSuccess;
    static size_t reserved = 0;
    const hipError_t e = reserve_lds((const void *)sym_fixed_kernel, lds, reserved);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sym_fixed_kernel, dim3(tiles), dim3(NT), lds, stream, P, off, cols);
    return hipGetLastError();
}

hipError_t launch_sym_count(const uint8_t *det, uint32_t det_row, const uint8_t *obs, uint32_t obs_row, uint64_t rows,
                            uint32_t *counts, hipStream_t stream) {
    if (!rows) return hipSuccess;
    hipLaunchKernelGGL(sym_count_kernel, dim3((uint32_t)((rows + 3) / 4)), dim3(256), 0, stream, det, det_row, obs,
                       obs_row, rows, counts);
    return hipGetLastError();
}

hipError_t launch_sym_fill(const uint8_t *det, uint32_t det_row, const uint8_t *obs, uint32_t obs_row, uint64_t rows,
                           uint32_t D, const uint32_t *offs, uint32_t *ent, hipStream_t stream) {
    if (!rows) return hipSuccess;
    hipLaunchKernelGGL(sym_fill_kernel, dim3((uint32_t)((rows + 3) / 4)), dim3(256), 0, stream, det, det_row, obs,
                       obs_row, rows, D, offs, ent);
    return hipGetLastError();
}

hipError_t launch_unpack_rows(const uint8_t *src, uint64_t src_stride, uint8_t *dst, uint64_t dst_stride, uint64_t rows,
                              uint32_t ncols, uint32_t base_char, hipStream_t stream) {
    const uint64_t total = rows * ncols;
    if (!total) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(unpack_rows_kernel, dim3(blocks), dim3(256), 0, stream, src, src_stride, dst, dst_stride, rows,
                       ncols, base_char);
    return hipGetLastError();
}

}  // namespace sa
