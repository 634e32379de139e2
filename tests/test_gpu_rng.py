"""Random streams on the GPU: the kernels' Philox4x32-10 against the CPU restatement (itself pinned on
the generator's published known answers), streams past 2^32 shot tiles, and the geometric gaps of
very rare mechanisms in groups of more than 10^9 positions."""

import numpy as np
import pytest

from oracle.philox import philox4x32_10

pytestmark = pytest.mark.gpu


def _device_philox(ctr, key):
    from stim_amd import _native as N
    ctr = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32).reshape(-1, 4))
    out = np.zeros_like(ctr)
    N.check(N.lib().mfs_philox(ctr.ctypes.data, len(ctr), int(key), out.ctypes.data))
    return out


def test_device_philox_known_answers():
    from test_rng_cpu import KAT
    for ctr, key, want in KAT:
        got = _device_philox([ctr], key[0] | (key[1] << 32))
        assert tuple(int(x) for x in got[0]) == want


def test_device_philox_matches_oracle_on_random_counters():
    rng = np.random.default_rng(5)
    ctr = rng.integers(0, 2**32, size=(1 << 16, 4), dtype=np.uint64).astype(np.uint32)
    for key in (0, 0xDEADBEEF124C, 2**64 - 1):
        want = philox4x32_10(ctr, (key & 0xFFFFFFFF, key >> 32))
        np.testing.assert_array_equal(_device_philox(ctr, key), want)


def _circuit():
    from stim_amd import gen
    return gen.generate('surface_code:rotated_memory_z', distance=5, rounds=5, after_clifford_depolarization=0.02,
                        after_reset_flip_probability=0.02, before_measure_flip_probability=0.02)


@pytest.mark.parametrize('engine', ['frame', 'symptom'])
def test_stream_past_2_32_tiles_does_not_repeat(engine):
    """Tile k and tile 2^32 + k draw different rows; setting the position back reproduces a call; the
    position advances by the tiles a call uses."""
    import stim_amd
    s = stim_amd.Circuit(_circuit()).compile_detector_sampler(seed=3)
    s.set_engine(engine)
    assert s.stream_position == 0
    shots = 4096
    a = s.sample(shots, bit_packed=True)
    used = s.stream_position
    assert used > 0
    fresh = stim_amd.Circuit(_circuit()).compile_detector_sampler(seed=3)
    fresh.set_engine(engine)
    np.testing.assert_array_equal(fresh.sample(shots, bit_packed=True), a)
    for k in (0, 5):
        s.stream_position = k
        base = s.sample(shots, bit_packed=True)
        s.stream_position = (1 << 32) + k
        far = s.sample(shots, bit_packed=True)
        assert s.stream_position == (1 << 32) + k + used
        assert far.any() and base.any()
        # Rows at the far position are not the rows of tile k (about 8% of rows are non-zero at this
        # noise level, so repeated tiles would match row for row).
        assert (far != base).any(axis=1).mean() > 0.5
        s.stream_position = (1 << 32) + k
        np.testing.assert_array_equal(s.sample(shots, bit_packed=True), far)
    # Far positions keep the distribution.
    s.stream_position = (7 << 40) + 11
    far = np.unpackbits(s.sample(1 << 16, bit_packed=True), axis=1, bitorder='little')
    s.stream_position = 11
    near = np.unpackbits(s.sample(1 << 16, bit_packed=True), axis=1, bitorder='little')
    a, b = far.sum(axis=0).astype(float), near.sum(axis=0).astype(float)
    z = (a - b) / np.sqrt(np.maximum(a + b, 1.0))
    assert np.abs(z).max() < 5.0


def test_batch_at_far_stream_position_equals_plain_call():
    import stim_amd
    from stim_amd.batch import BatchSampler
    text = _circuit()
    s = stim_amd.Circuit(text).compile_detector_sampler(seed=8)
    ref = stim_amd.Circuit(text).compile_detector_sampler(seed=8)
    for x in (s, ref):
        x.engine_info()
        x.stream_position = (3 << 32) + 17
    got = BatchSampler([(s, 5000)]).run()
    d, o = ref.sample(5000, bit_packed=True, separate_observables=True)
    np.testing.assert_array_equal(got[0][0].cpu().numpy(), d)
    np.testing.assert_array_equal(got[0][1].cpu().numpy(), o)
    assert s.stream_position == ref.stream_position


def test_rare_mechanisms_in_group_above_1e9_positions():
    """250,000 mechanisms of p = 5e-10 on one detector: one group of 250,000 x 8,192 > 2^30 positions
    per tile, where most gaps exceed 10^9. The detector's rate must match 1 - (1 - p)^250000."""
    import stim_amd
    n, p = 250000, 5e-10
    s = stim_amd.DetectorErrorModel(f'repeat {n} {{\n    error({p}) D0\n}}\n').compile_sampler(seed=2)
    shots = 1 << 22
    d, _, _ = s.sample(shots, bit_packed=True)
    count = int(np.unpackbits(d[:, 0:1], axis=1, bitorder='little')[:, 0].sum())
    rate = 1.0 - (1.0 - p) ** n
    mu = shots * rate
    assert abs(count - mu) < 5.0 * np.sqrt(mu * (1 - rate)), (count, mu)
