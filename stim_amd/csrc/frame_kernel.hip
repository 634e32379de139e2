// Frame-simulation kernel for gfx950 (MI355X).
//
// One workgroup owns a tile of 32*T shots. Its whole Pauli frame (X and Z rows, one T-word row per
// qubit, shots along the bits), the measurement-record ring, 256 staged detector rows and the
// observable rows live in LDS for the entire circuit. The workgroup walks the device program
// (program.h) in order; each op spreads its items (targets, pairs, detectors, noise chunks) over
// the workgroup's lanes, and ops that touch a row an earlier unsynchronized op touched are separated by a
// workgroup barrier (placed by the host compiler).
//
// Reference semantics restated here:
//   frame rules            src/stim/simulators/frame_simulator.inl:172-630
//   sparse noise           RareErrorIterator (src/stim/util_bot/probability_util.h:33-71): a Bernoulli(p)
//                          process over (site, shot) positions sampled by geometric gaps; here the
//                          positions of a tile are cut into chunks, each chunk an independent
//                          Philox4x32-10 stream (Salmon et al., SC'11), which keeps the process exact
//                          because geometric gaps are memoryless.
//   detectors/observables  frame_simulator.inl:221-252 (XOR of record rows / frame rows)
//   shot-major output      src/stim/py/numpy.pybind.cc:20-63 (bit m of shot s at byte m/8, bit m%8);
//                          the transpose is done 32x32 bits at a time across lanes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bitops.h"
#include "launch_util.h"
#include "program.h"
#include "rng.h"

using namespace sa;

namespace {

// Workgroup widths: 256 lanes, or 1024 for circuits whose tile fills most of a CU's LDS (one
// workgroup per CU), so that the CU still has 16 waves to run.
constexpr int NT_SMALL = 256;
constexpr int NT_LARGE = 1024;

template <int T>
struct alignas(4 * T) Wd {
    uint32_t v[T];
};

template <int T>
__device__ __forceinline__ Wd<T> wxor(Wd<T> a, const Wd<T> &b) {
#pragma unroll
    for (int k = 0; k < T; k++) a.v[k] ^= b.v[k];
    return a;
}
template <int T>
__device__ __forceinline__ Wd<T> wzero() {
    Wd<T> r;
#pragma unroll
    for (int k = 0; k < T; k++) r.v[k] = 0;
    return r;
}
template <int T>
__device__ __forceinline__ Wd<T> &row(uint32_t *base, uint32_t r) {
    return *reinterpret_cast<Wd<T> *>(base + (size_t)r * T);
}

// Stream tags keep the counter spaces of different uses apart.
constexpr uint32_t TAG_NOISE = 0u << 28;
constexpr uint32_t TAG_RAND = 1u << 28;
constexpr uint32_t TAG_INIT = 2u << 28;
constexpr uint32_t TAG_DENSE = 3u << 28;

struct Rng {
    uint32_t k0, k1;
    uint32_t tile;
    __device__ uint4 draw(uint32_t a, uint32_t b, uint32_t tag_and_id) const {
        return philox10(make_uint4(a, b, tile, tag_and_id), k0, k1);
    }
};

// Geometric gap: number of failures before the next success of a Bernoulli(p) process,
// floor(log(u) / log(1 - p)) with u uniform in (0, 1].
__device__ __forceinline__ uint64_t geometric_gap(uint32_t r, float inv_log2_1mp) {
    float u = ((float)(r >> 8) + 1.0f) * (1.0f / 16777216.0f);
    float g = __log2f(u) * inv_log2_1mp;
    if (!(g < 1.0e12f)) return (uint64_t)1 << 40;
    return (uint64_t)g;
}

__device__ __forceinline__ uint32_t pick_outcome(const uint32_t *otab, uint32_t woff, uint32_t nout, uint32_t r) {
    for (uint32_t i = 0; i + 1 < nout; i++) {
        if (r <= otab[woff + 2 * i]) return otab[woff + 2 * i + 1];
    }
    return otab[woff + 2 * (nout - 1) + 1];
}

// Bernoulli(p) bits for 32 shots by comparing 24-bit uniforms bit-sliced across the word.
__device__ __forceinline__ uint32_t bernoulli_word(float p, const Rng &rng, uint32_t a, uint32_t id) {
    if (p >= 1.0f) return 0xFFFFFFFFu;
    uint32_t P = (uint32_t)(p * 16777216.0f);
    uint32_t lt = 0, eq = 0xFFFFFFFFu;
    for (int blk = 0; blk < 6; blk++) {
        uint4 r = rng.draw(a, (uint32_t)blk, TAG_DENSE | id);
        uint32_t rs[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int bit = 23 - (blk * 4 + k);
            uint32_t pb = (P >> bit) & 1u;
            if (pb) {
                lt |= eq & ~rs[k];
                eq &= rs[k];
            } else {
                eq &= ~rs[k];
            }
        }
    }
    return lt;
}

template <int T>
struct Tile {
    uint32_t *X, *Z, *REC, *STG, *OBS, *CHAIN;
};

// Applies one noise outcome to shot s of the tile. site is the site within op.
template <int T>
__device__ __forceinline__ void apply_hit(const Tile<T> &t, const DevOp &op, const uint32_t *__restrict__ tgt,
                                          uint32_t site, uint32_t s, uint32_t code) {
    const uint32_t w = s >> 5;
    const uint32_t bit = 1u << (s & 31);
    switch (op.code & 0xFF) {
        case OP_NOISE1: {
            uint32_t q = tgt[op.toff + site];
            if (code & 1) atomicXor(&t.X[q * T + w], bit);
            if (code & 2) atomicXor(&t.Z[q * T + w], bit);
            break;
        }
        case OP_NOISE2: {
            uint32_t q1 = tgt[op.toff + 2 * site], q2 = tgt[op.toff + 2 * site + 1];
            if (code & 1) atomicXor(&t.X[q1 * T + w], bit);
            if (code & 2) atomicXor(&t.Z[q1 * T + w], bit);
            if (code & 4) atomicXor(&t.X[q2 * T + w], bit);
            if (code & 8) atomicXor(&t.Z[q2 * T + w], bit);
            break;
        }
        case OP_MFLIP:
            if (code & 1) atomicXor(&t.REC[(op.base + site) * T + w], bit);
            break;
        case OP_CORR: {
            if (!(code & 1)) break;
            uint32_t old = atomicOr(&t.CHAIN[w], bit);
            if (old & bit) break;
            for (uint32_t i = 0; i < op.base; i++) {
                uint32_t term = tgt[op.toff + i];
                uint32_t q = term & TERM_QMASK;
                if (term & TERM_X) atomicXor(&t.X[q * T + w], bit);
                if (term & TERM_Z) atomicXor(&t.Z[q * T + w], bit);
            }
            break;
        }
        case OP_HERALD: {
            uint32_t q = tgt[op.toff + site];
            if (code & 4) atomicOr(&t.REC[(op.base + site) * T + w], bit);
            if (code & 1) atomicXor(&t.X[q * T + w], bit);
            if (code & 2) atomicXor(&t.Z[q * T + w], bit);
            break;
        }
        default:
            break;
    }
}

// Writes staged rows [0, n) (entries base .. base + n) of the tile to one output.
template <int T, int NT>
__device__ void flush_rows(const uint32_t *rows, uint32_t n, uint32_t base, const DevOut &out, uint64_t shot0,
                           uint64_t shots_valid) {
    const uint32_t tid = threadIdx.x;
    if (out.packed) {
        // Blocks of 32 rows x 32 shots; each half-wave transposes one block.
        const uint32_t nblk = ((n + 31) / 32) * T;
        const uint32_t half = tid >> 5, lane = tid & 31;
        for (uint32_t b = half; b < ((nblk + NT / 32 - 1) / (NT / 32)) * (NT / 32); b += NT / 32) {
            // All lanes of a half-wave take part in the shuffles; blocks past the end are masked.
            const bool live = b < nblk;
            const uint32_t rb = live ? b / T : 0, w = live ? b % T : 0;
            const uint32_t r = rb * 32 + lane;
            uint32_t x = (live && r < n) ? rows[r * T + w] : 0u;
            x = transpose32(x, lane);
            const uint64_t shot = shot0 + w * 32 + lane;
            if (live && shot < shots_valid) {
                uint8_t *p = out.ptr + shot * out.stride + (out.col0 >> 3) + ((base + rb * 32) >> 3);
                const uint32_t remaining = n - rb * 32;
                const uint32_t nbytes = remaining >= 32 ? 4 : (remaining + 7) / 8;
                store_bytes(p, x, nbytes);
            }
        }
    } else {
        const uint32_t W = 32 * T;
        for (uint32_t i = tid; i < n * W; i += NT) {
            const uint32_t s = i / n, r = i % n;
            const uint64_t shot = shot0 + s;
            if (shot < shots_valid) {
                uint32_t bitv = (rows[r * T + (s >> 5)] >> (s & 31)) & 1u;
                out.ptr[shot * out.stride + out.col0 + base + r] = (uint8_t)(bitv + out.base_char);
            }
        }
    }
}

template <int T>
__device__ __forceinline__ Wd<T> rand_words(const Rng &rng, uint32_t op_i, uint32_t item) {
    uint4 r = rng.draw(item, op_i, TAG_RAND);
    Wd<T> out;
    const uint32_t rs[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int k = 0; k < T; k++) out.v[k] = rs[k];
    return out;
}

template <int T, int NT>
__global__ __launch_bounds__(NT) void frame_sim_kernel(const uint32_t *__restrict__ blob,
                                                       const uint32_t *__restrict__ ntgt,
                                                       const uint32_t *__restrict__ otab, LaunchParams P) {
    extern __shared__ uint32_t lds[];
    // LDS layout: two blob staging buffers, then the tile (frame rows, record ring, staged
    // detector rows, observable rows, chain mask).
    uint32_t *const STAGE = lds;
    Tile<T> t;
    t.X = lds + 2 * STAGE_WORDS;
    t.Z = t.X + P.num_qubits * T;
    t.REC = t.Z + P.num_qubits * T;
    t.STG = t.REC + P.ring * T;
    t.OBS = t.STG + STAGE_ROWS * T;
    t.CHAIN = t.OBS + P.num_obs * T;
    const uint32_t tid = threadIdx.x;
    const uint32_t W = 32 * T;
    const uint64_t shot0 = (uint64_t)blockIdx.x * W;
    const uint64_t tile64 = P.tile0 + blockIdx.x;
    Rng rng{(uint32_t)P.seed, frame_key_hi((uint32_t)(P.seed >> 32), tile64), (uint32_t)tile64};
    const bool randomize = P.randomize != 0;

    for (uint32_t i = tid; i < P.num_qubits; i += NT) {
        row<T>(t.X, i) = wzero<T>();
        row<T>(t.Z, i) = randomize ? rand_words<T>(rng, 0xFFFFFFFFu, i) : wzero<T>();
    }
    for (uint32_t i = tid; i < (P.num_obs + 1) * T; i += NT) t.OBS[i] = 0;
    for (uint32_t i = tid; i < P.blob0_len; i += NT) STAGE[i] = blob[i];
    __syncthreads();

    uint32_t fptr = 0, fend = 0;
    if (P.mode == MODE_FIXED) {
        fptr = P.fault_off[blockIdx.x];
        fend = P.fault_off[blockIdx.x + 1];
    }

    uint32_t buf = 0;
    while (P.blob0_len) {
        const uint32_t *cur = STAGE + buf * STAGE_WORDS;
        const uint32_t n_ops = cur[0], nbeg = cur[1], nlen = cur[2], op0 = cur[3];
        // Prefetch the next blob into registers; it lands in the other staging buffer after this
        // epoch's ops, so its latency hides behind this epoch's work.
        uint32_t pre[STAGE_WORDS / NT];
#pragma unroll
        for (uint32_t k = 0; k < STAGE_WORDS / NT; k++) {
            const uint32_t idx = tid + k * NT;
            pre[k] = idx < nlen ? blob[nbeg + idx] : 0u;
        }
        const DevOp *hdrs = reinterpret_cast<const DevOp *>(cur + BLOB_HEADER);

        for (uint32_t j = 0; j < n_ops; j++) {
            const DevOp op = hdrs[j];
            const uint32_t oi = op0 + j;
            const uint32_t code = op.code & 0xFF;
            const bool rnd = randomize && (op.code & OPF_RANDOMIZE);
            const uint32_t *w = cur + op.toff;
            switch (code) {
                case OP_H:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t q = w[i];
                        Wd<T> x = row<T>(t.X, q), z = row<T>(t.Z, q);
                        row<T>(t.X, q) = z;
                        row<T>(t.Z, q) = x;
                    }
                    break;
                case OP_HXY:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t q = w[i];
                        row<T>(t.Z, q) = wxor(row<T>(t.Z, q), row<T>(t.X, q));
                    }
                    break;
                case OP_HYZ:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t q = w[i];
                        row<T>(t.X, q) = wxor(row<T>(t.X, q), row<T>(t.Z, q));
                    }
                    break;
                case OP_CXYZ:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t q = w[i];
                        Wd<T> x = row<T>(t.X, q), z = row<T>(t.Z, q);
                        x = wxor(x, z);
                        z = wxor(z, x);
                        row<T>(t.X, q) = x;
                        row<T>(t.Z, q) = z;
                    }
                    break;
                case OP_CZYX:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t q = w[i];
                        Wd<T> x = row<T>(t.X, q), z = row<T>(t.Z, q);
                        z = wxor(z, x);
                        x = wxor(x, z);
                        row<T>(t.X, q) = x;
                        row<T>(t.Z, q) = z;
                    }
                    break;
                case OP_CX: case OP_CY: case OP_CZ: case OP_SWAP: case OP_ISWAP: case OP_CXSWAP: case OP_CZSWAP:
                case OP_SWAPCX: case OP_SQRTXX: case OP_SQRTYY: case OP_SQRTZZ: case OP_XCX: case OP_XCY:
                case OP_YCX: case OP_YCY:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t qa = w[2 * i], qb = w[2 * i + 1];
                        Wd<T> x1 = row<T>(t.X, qa), z1 = row<T>(t.Z, qa);
                        Wd<T> x2 = row<T>(t.X, qb), z2 = row<T>(t.Z, qb);
                        switch (code) {
                            case OP_CX: z1 = wxor(z1, z2); x2 = wxor(x2, x1); break;
                            case OP_CY: z1 = wxor(z1, wxor(x2, z2)); z2 = wxor(z2, x1); x2 = wxor(x2, x1); break;
                            case OP_CZ: z1 = wxor(z1, x2); z2 = wxor(z2, x1); break;
                            case OP_SWAP: { Wd<T> a = x1; x1 = x2; x2 = a; a = z1; z1 = z2; z2 = a; } break;
                            case OP_ISWAP: {
                                Wd<T> dx = wxor(x1, x2);
                                Wd<T> t1 = wxor(z1, dx), t2 = wxor(z2, dx);
                                z1 = t2; z2 = t1;
                                Wd<T> a = x1; x1 = x2; x2 = a;
                            } break;
                            case OP_CXSWAP: z2 = wxor(z2, z1); z1 = wxor(z1, z2); x1 = wxor(x1, x2); x2 = wxor(x2, x1); break;
                            case OP_CZSWAP: {
                                Wd<T> a = x1; x1 = x2; x2 = a; a = z1; z1 = z2; z2 = a;
                                z1 = wxor(z1, x2); z2 = wxor(z2, x1);
                            } break;
                            case OP_SWAPCX: z1 = wxor(z1, z2); z2 = wxor(z2, z1); x2 = wxor(x2, x1); x1 = wxor(x1, x2); break;
                            case OP_SQRTXX: { Wd<T> d = wxor(z1, z2); x1 = wxor(x1, d); x2 = wxor(x2, d); } break;
                            case OP_SQRTYY: {
                                Wd<T> d = wxor(wxor(x1, z1), wxor(x2, z2));
                                x1 = wxor(x1, d); z1 = wxor(z1, d); x2 = wxor(x2, d); z2 = wxor(z2, d);
                            } break;
                            case OP_SQRTZZ: { Wd<T> d = wxor(x1, x2); z1 = wxor(z1, d); z2 = wxor(z2, d); } break;
                            case OP_XCX: x1 = wxor(x1, z2); x2 = wxor(x2, z1); break;
                            case OP_XCY: x1 = wxor(x1, wxor(x2, z2)); x2 = wxor(x2, z1); z2 = wxor(z2, z1); break;
                            case OP_YCX: x2 = wxor(x2, wxor(x1, z1)); x1 = wxor(x1, z2); z1 = wxor(z1, z2); break;
                            case OP_YCY: {
                                Wd<T> y1 = wxor(x1, z1), y2 = wxor(x2, z2);
                                x1 = wxor(x1, y2); z1 = wxor(z1, y2); x2 = wxor(x2, y1); z2 = wxor(z2, y1);
                            } break;
                            default: break;
                        }
                        row<T>(t.X, qa) = x1;
                        row<T>(t.Z, qa) = z1;
                        row<T>(t.X, qb) = x2;
                        row<T>(t.Z, qb) = z2;
                    }
                    break;
                case OP_FEEDBACK:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t slot = w[3 * i], q = w[3 * i + 1], mask = w[3 * i + 2];
                        const Wd<T> r = row<T>(t.REC, slot);
                        if (mask & 1) row<T>(t.X, q) = wxor(row<T>(t.X, q), r);
                        if (mask & 2) row<T>(t.Z, q) = wxor(row<T>(t.Z, q), r);
                    }
                    break;
                case OP_R: case OP_RX: case OP_RY:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t q = w[i];
                        const Wd<T> r = rnd ? rand_words<T>(rng, oi, i) : wzero<T>();
                        if (code == OP_R) {
                            row<T>(t.X, q) = wzero<T>();
                            if (rnd) row<T>(t.Z, q) = wxor(row<T>(t.Z, q), r);
                        } else if (code == OP_RX) {
                            row<T>(t.Z, q) = wzero<T>();
                            if (rnd) row<T>(t.X, q) = wxor(row<T>(t.X, q), r);
                        } else {
                            Wd<T> z = rnd ? wxor(row<T>(t.Z, q), r) : row<T>(t.Z, q);
                            row<T>(t.Z, q) = z;
                            row<T>(t.X, q) = z;
                        }
                    }
                    break;
                case OP_M: case OP_MX: case OP_MY: case OP_MR: case OP_MRX: case OP_MRY:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t q = w[i];
                        Wd<T> x = row<T>(t.X, q), z = row<T>(t.Z, q);
                        const Wd<T> r = rnd ? rand_words<T>(rng, oi, i) : wzero<T>();
                        Wd<T> rec;
                        if (code == OP_M || code == OP_MR) {
                            rec = x;
                            if (code == OP_MR) x = wzero<T>();
                            z = wxor(z, r);
                        } else if (code == OP_MX || code == OP_MRX) {
                            rec = z;
                            if (code == OP_MRX) z = wzero<T>();
                            x = wxor(x, r);
                        } else {
                            rec = wxor(x, z);
                            if (code == OP_MY) {
                                x = wxor(x, r);
                                z = wxor(z, r);
                            } else {
                                z = wxor(z, r);
                                x = z;
                            }
                        }
                        row<T>(t.X, q) = x;
                        row<T>(t.Z, q) = z;
                        row<T>(t.REC, op.base + i) = rec;
                    }
                    break;
                case OP_MPP:
                case OP_SPP:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t b = w[i], e2 = w[i + 1];
                        Wd<T> v = wzero<T>();
                        for (uint32_t k = b; k < e2; k++) {
                            const uint32_t term = w[k];
                            if (term & TERM_PAULI) {
                                const uint32_t q = term & TERM_QMASK;
                                if (term & TERM_X) v = wxor(v, row<T>(t.Z, q));
                                if (term & TERM_Z) v = wxor(v, row<T>(t.X, q));
                            } else {
                                v = wxor(v, row<T>(t.REC, term));
                            }
                        }
                        Wd<T> apply;
                        if (code == OP_MPP) {
                            row<T>(t.REC, op.base + i) = v;
                            if (!rnd) continue;
                            apply = rand_words<T>(rng, oi, i);
                        } else {
                            apply = v;
                        }
                        for (uint32_t k = b; k < e2; k++) {
                            const uint32_t term = w[k];
                            if (!(term & TERM_PAULI)) continue;
                            const uint32_t q = term & TERM_QMASK;
                            if (term & TERM_X) row<T>(t.X, q) = wxor(row<T>(t.X, q), apply);
                            if (term & TERM_Z) row<T>(t.Z, q) = wxor(row<T>(t.Z, q), apply);
                        }
                    }
                    break;
                case OP_RECZERO:
                    for (uint32_t i = tid; i < op.n; i += NT) row<T>(t.REC, op.base + i) = wzero<T>();
                    break;
                case OP_DET:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t b = w[i], e2 = w[i + 1];
                        Wd<T> v = wzero<T>();
                        for (uint32_t k = b; k < e2; k++) v = wxor(v, row<T>(t.REC, w[k]));
                        if (P.det_invert) {
                            const uint32_t d = op.base + i;
                            if ((P.det_invert[d >> 5] >> (d & 31)) & 1u) {
#pragma unroll
                                for (int k = 0; k < T; k++) v.v[k] = ~v.v[k];
                            }
                        }
                        row<T>(t.STG, (op.base + i) % STAGE_ROWS) = v;
                    }
                    break;
                case OP_OBS:
                    for (uint32_t i = tid; i < op.n; i += NT) {
                        const uint32_t b = w[i], e2 = w[i + 1];
                        const uint32_t idx = w[b];
                        Wd<T> v = wzero<T>();
                        for (uint32_t k = b + 1; k < e2; k++) {
                            const uint32_t term = w[k];
                            if (term & TERM_PAULI) {
                                const uint32_t q = term & TERM_QMASK;
                                if (term & TERM_X) v = wxor(v, row<T>(t.Z, q));
                                if (term & TERM_Z) v = wxor(v, row<T>(t.X, q));
                            } else {
                                v = wxor(v, row<T>(t.REC, term));
                            }
                        }
#pragma unroll
                        for (int k = 0; k < T; k++) {
                            if (v.v[k]) atomicXor(&t.OBS[idx * T + k], v.v[k]);
                        }
                    }
                    break;
                case OP_FLUSH:
                    if (P.dets.enabled) flush_rows<T, NT>(t.STG, op.n, op.base, P.dets, shot0, P.shots);
                    break;
                case OP_FLUSH_OBS:
                    if (P.obs.enabled) flush_rows<T, NT>(t.OBS, op.n, 0, P.obs, shot0, P.shots);
                    break;
                case OP_CHAINCLR:
                    if (tid < T) t.CHAIN[tid] = 0;
                    break;
                case OP_NOISE1: case OP_NOISE2: case OP_MFLIP: case OP_CORR: case OP_HERALD: {
                    if (P.mode == MODE_FIXED) {
                        // Faults of this tile are sorted by op; consume the prefix that belongs to this op.
                        while (true) {
                            const uint32_t idx = fptr + tid;
                            bool mine = false;
                            if (idx < fend) {
                                const DevFault f = P.faults[idx];
                                if (f.op == oi) {
                                    mine = true;
                                    apply_hit<T>(t, op, ntgt, f.site, f.shot, f.code);
                                }
                            }
                            const uint32_t cnt = __syncthreads_count(mine);
                            fptr += cnt;
                            if (cnt < NT) break;
                        }
                        break;
                    }
                    if (P.mode != MODE_SAMPLED || !(op.p > 0.0f)) break;
                    const uint32_t nout = op.nout & 0xFFFF;
                    if (op.code & OPF_DENSE) {
                        // Dense draw: Bernoulli word per (site, word), then one outcome per fired shot.
                        for (uint32_t i = tid; i < op.n * T; i += NT) {
                            const uint32_t site = i / T, wd = i % T;
                            uint32_t fire = bernoulli_word(op.p, rng, i, oi);
                            uint32_t ctr = 0;
                            uint4 r = make_uint4(0, 0, 0, 0);
                            while (fire) {
                                const uint32_t b = __ffs(fire) - 1;
                                fire &= fire - 1;
                                if ((ctr & 3) == 0) r = rng.draw(i, 0x40000000u | ctr, TAG_NOISE | oi);
                                const uint32_t rs = (ctr & 3) == 0 ? r.x : (ctr & 3) == 1 ? r.y : (ctr & 3) == 2 ? r.z : r.w;
                                ctr++;
                                apply_hit<T>(t, op, ntgt, site, wd * 32 + b, pick_outcome(otab, op.woff, nout, rs));
                            }
                        }
                        break;
                    }
                    // The tile's (site, shot) positions are cut into chunks spread over all lanes (at
                    // most ~4/p positions each); each chunk walks its own geometric gaps, two hits per
                    // Philox block (gap, outcome, gap, outcome).
                    const uint64_t N = (uint64_t)op.n * W;
                    const uint32_t cmax = (op.nout >> 16) & 0xFF;
                    uint32_t clog = 6;
                    while (clog < cmax + 2 && ((uint64_t)NT << clog) < N) clog++;
                    const uint64_t C = (uint64_t)1 << clog;
                    const uint64_t nch = (N + C - 1) >> clog;
                    for (uint64_t ch = tid; ch < nch; ch += NT) {
                        const uint64_t lo = ch << clog;
                        const uint64_t hi = lo + C < N ? lo + C : N;
                        uint64_t pos = lo;
                        uint32_t ctr = 0;
                        while (true) {
                            const uint4 r = rng.draw(ctr++, (uint32_t)ch, TAG_NOISE | oi);
                            pos += geometric_gap(r.x, op.inv_log);
                            if (pos >= hi) break;
                            apply_hit<T>(t, op, ntgt, (uint32_t)(pos / W), (uint32_t)(pos % W),
                                         pick_outcome(otab, op.woff, nout, r.y));
                            pos += 1 + geometric_gap(r.z, op.inv_log);
                            if (pos >= hi) break;
                            apply_hit<T>(t, op, ntgt, (uint32_t)(pos / W), (uint32_t)(pos % W),
                                         pick_outcome(otab, op.woff, nout, r.w));
                            pos++;
                        }
                    }
                    break;
                }
                default:
                    break;
            }
        }
        if (nlen == 0) break;
        uint32_t *nst = STAGE + (buf ^ 1) * STAGE_WORDS;
#pragma unroll
        for (uint32_t k = 0; k < STAGE_WORDS / NT; k++) {
            const uint32_t idx = tid + k * NT;
            if (idx < nlen) nst[idx] = pre[k];
        }
        __syncthreads();
        buf ^= 1;
    }
}

__global__ void pack_rows_kernel(const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride,
                                 uint64_t rows, uint32_t ncols) {
    const uint32_t nb = (ncols + 7) / 8;
    const uint64_t total = rows * nb;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = i / nb;
        const uint32_t c = (uint32_t)(i % nb);
        uint32_t v = 0;
        for (uint32_t b = 0; b < 8; b++) {
            const uint32_t col = 8 * c + b;
            if (col < ncols) v |= (uint32_t)(src[r * src_stride + col] & 1u) << b;
        }
        dst[r * dst_stride + c] = (uint8_t)v;
    }
}

}  // namespace

// Host-side launcher (C++ linkage, called from sampler.cc).
namespace sa {

hipError_t launch_frame_kernel(int T, const DevProgram &prog, const LaunchParams &P, uint32_t num_tiles, size_t lds,
                               hipStream_t stream) {
    if (num_tiles == 0) return hipSuccess;
    // A tile of more than 64 KB leaves room for at most two workgroups per CU: use the wide workgroup.
    const bool wide = lds > 64 * 1024;
    dim3 grid(num_tiles), block(wide ? NT_LARGE : NT_SMALL);
#define STIM_AMD_LAUNCH(TW, NB)                                                                                 \
    {                                                                                                           \
        static size_t reserved = 0;                                                                             \
        const hipError_t e = reserve_lds((const void *)frame_sim_kernel<TW, NB>, lds, reserved);                \
        if (e != hipSuccess) return e;                                                                          \
    }                                                                                                           \
    hipLaunchKernelGGL((frame_sim_kernel<TW, NB>), grid, block, lds, stream, prog.blob, prog.ntgt, prog.otab, P);
    switch (T * 2 + (wide ? 1 : 0)) {
        case 2: STIM_AMD_LAUNCH(1, NT_SMALL) break;
        case 3: STIM_AMD_LAUNCH(1, NT_LARGE) break;
        case 4: STIM_AMD_LAUNCH(2, NT_SMALL) break;
        case 5: STIM_AMD_LAUNCH(2, NT_LARGE) break;
        case 8: STIM_AMD_LAUNCH(4, NT_SMALL) break;
        case 9: STIM_AMD_LAUNCH(4, NT_LARGE) break;
        default: return hipErrorInvalidValue;
    }
#undef STIM_AMD_LAUNCH
    return hipGetLastError();
}

hipError_t launch_pack_rows(const uint8_t *src, size_t src_stride, uint8_t *dst, size_t dst_stride, uint64_t rows,
                            uint32_t ncols, hipStream_t stream) {
    const uint64_t total = rows * ((ncols + 7) / 8);
    if (total == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(pack_rows_kernel, dim3(blocks), dim3(256), 0, stream, src, src_stride, dst, dst_stride, rows, ncols);
    return hipGetLastError();
}

}  // namespace sa
