"""CPU checks of the random streams: the Philox restatement against the published known answers, and
the counters the kernels build from 64-bit tile indices (mfs_rng_counter, host code shared with the
kernels), including tiles past 2^32."""

import ctypes

import numpy as np
import pytest

from oracle.philox import philox4x32_10

# Known-answer vectors of Philox4x32-10 published with the generator (Salmon et al., SC'11):
# (counter, key) -> output.
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize('ctr,key,want', KAT)
def test_philox_oracle_known_answers(ctr, key, want):
    assert tuple(int(x) for x in philox4x32_10(ctr, key)[0]) == want


def test_philox_oracle_planted_mistake_fails():
    # One round fewer (a dropped round) must miss the known answers.
    import oracle.philox as ph
    c = np.zeros((1, 4), np.uint64)
    k = np.zeros((1, 2), np.uint64)
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(9):
        p0 = np.uint64(ph.M0) * c[:, 0]
        p1 = np.uint64(ph.M1) * c[:, 2]
        c = np.stack([(p1 >> np.uint64(32)) ^ c[:, 1] ^ k[:, 0], p1 & mask,
                      (p0 >> np.uint64(32)) ^ c[:, 3] ^ k[:, 1], p0 & mask], axis=1)
        k[:, 0] = (k[:, 0] + np.uint64(ph.W0)) & mask
        k[:, 1] = (k[:, 1] + np.uint64(ph.W1)) & mask
    assert tuple(int(x) for x in c[0]) != KAT[0][2]


def _counter(kernel, key, tile, lane, draw, tag):
    from stim_amd import _native as N
    out = (ctypes.c_uint32 * 6)()
    N.check(N.lib().mfs_rng_counter(kernel, key, tile, lane, draw, tag, out))
    return tuple(out)


KEY = 0x0123456789ABCDEF


@pytest.mark.parametrize('kernel', [0, 1])
def test_counters_below_2_32_tiles_are_the_plain_layout(kernel):
    # Tiles below 2^32 keep (draw, lane, tile, tag) under the sampler key, so earlier outputs hold.
    for tile in (0, 1, 12345, 2**32 - 1):
        for lane in (0, 7, 255):
            assert _counter(kernel, KEY, tile, lane, 3, 0x60000002) == (
                3, lane, tile & 0xFFFFFFFF, 0x60000002, KEY & 0xFFFFFFFF, KEY >> 32)


@pytest.mark.parametrize('kernel', [0, 1])
def test_tiles_past_2_32_get_new_streams(kernel):
    for k in (0, 1, 999):
        for hi in (1, 2, 255, 2**20 + 3):
            a = _counter(kernel, KEY, k, 5, 0, 0)
            b = _counter(kernel, KEY, (hi << 32) + k, 5, 0, 0)
            assert a != b
            pa = philox4x32_10(a[:4], a[4:])
            pb = philox4x32_10(b[:4], b[4:])
            assert not np.array_equal(pa, pb)


@pytest.mark.parametrize('kernel', [0, 1])
def test_counter_and_key_distinct_over_tiles_and_lanes(kernel):
    seen = set()
    for hi in range(64):
        for lo in (0, 1, 77):
            for lane in range(0, 256, 17):
                c = _counter(kernel, KEY, (hi << 32) | lo, lane, 0, 0)
                assert c not in seen
                seen.add(c)


def test_symptom_lane_limit():
    from stim_amd import _native as N
    out = (ctypes.c_uint32 * 6)()
    assert N.lib().mfs_rng_counter(0, KEY, 0, 256, 0, 0, out) == N.MFS_ERR_INVALID_ARGUMENT


def test_library_build_id_matches_tree():
    from stim_amd import _native as N
    from stim_amd._buildid import source_hash
    assert N.build_id() == source_hash()
    assert N.check_build_id() == source_hash()
