"""Philox4x32-10 on the CPU (test infrastructure; checker only, never on the product path).

Restates the generator of Salmon, Moraes, Dror and Shaw, "Parallel Random Numbers: As Easy as 1, 2,
3" (SC'11), which the GPU samplers draw from (the reference itself uses std::mt19937_64,
src/stim/util_bot/probability_util.h:40-76, whose stream no reference test pins). Pinned by the
paper's published known-answer vectors for 10 rounds (tests/test_rng_cpu.py); the GPU generator is
checked against this restatement on random counters (tests/test_gpu_rng.py).

Round: (hi(M0*x0), lo(M0*x0)) and (hi(M1*x2), lo(M1*x2)) give
x' = (hi1 ^ x1 ^ k0, lo1, hi0 ^ x3 ^ k1, lo0); the key is bumped by the Weyl constants between rounds.
"""

import numpy as np

M0 = 0xD2511F53
M1 = 0xCD9E8D57
W0 = 0x9E3779B9
W1 = 0xBB67AE85


def philox4x32_10(ctr, key):
    """ctr: (n, 4) uint32 counters; key: (k0, k1) or (n, 2) uint32. Returns (n, 4) uint32."""
    c = np.asarray(ctr, dtype=np.uint64).reshape(-1, 4).copy()
    k = np.asarray(key, dtype=np.uint64)
    k = np.broadcast_to(k.reshape(-1, 2), (c.shape[0], 2)).copy()
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(M0) * c[:, 0]
        p1 = np.uint64(M1) * c[:, 2]
        x0 = (p1 >> np.uint64(32)) ^ c[:, 1] ^ k[:, 0]
        x1 = p1 & mask
        x2 = (p0 >> np.uint64(32)) ^ c[:, 3] ^ k[:, 1]
        x3 = p0 & mask
        c = np.stack([x0, x1, x2, x3], axis=1)
        k[:, 0] = (k[:, 0] + np.uint64(W0)) & mask
        k[:, 1] = (k[:, 1] + np.uint64(W1)) & mask
    return c.astype(np.uint32)
