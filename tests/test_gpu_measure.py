"""Measurement sampler on the GPU (stim_amd.Circuit.compile_sampler()).

Pinned by: the reference CLI's seeded `stim sample` counts per measurement (meas_counts.json; a
two-sample binomial test, Bonferroni over every measurement), the reference sample (itself pinned
bit-exact against the reference CLI) for every deterministic result of noiseless circuits, and the
parity of every detector of a noiseless circuit (its measurements XOR their reference bits is 0).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gen(code, d, r, p):
    from stim_amd import gen
    return gen.generate(code, distance=d, rounds=r, after_clifford_depolarization=p, after_reset_flip_probability=p,
                        before_measure_flip_probability=p)


def _zcrit(n, alpha=1e-4):
    from scipy.stats import norm
    return abs(norm.ppf(alpha / (2 * n)))


def _bits(a, m):
    return np.unpackbits(a, axis=1, bitorder='little')[:, :m].astype(bool)


def test_counts_match_reference_cli(golden):
    import stim_amd
    fixtures = golden.load('meas_counts.json')
    zs = []
    for name, f in fixtures.items():
        text = f['circuit'] or golden.circuit(name)
        s = stim_amd.Circuit(text).compile_sampler(seed=11)
        shots = 1 << 20
        m = s.num_measurements
        assert m == len(f['counts'])
        ours = _bits(s.sample(shots, bit_packed=True), m).sum(axis=0)
        theirs = np.asarray(f['counts'], dtype=float)
        n2 = f['shots']
        for k in range(m):
            p = (ours[k] + theirs[k]) / (shots + n2)
            if p in (0.0, 1.0):
                continue
            z = (ours[k] / shots - theirs[k] / n2) / np.sqrt(p * (1 - p) * (1 / shots + 1 / n2))
            zs.append((abs(z), name, k))
    worst = max(zs)
    assert worst[0] < _zcrit(len(zs)), worst


def test_noiseless_results_follow_the_reference_sample():
    import stim_amd
    from oracle import stim_text as st
    text = _gen('surface_code:rotated_memory_x', 5, 4, 0.0)
    c = stim_amd.Circuit(text)
    ref = c.reference_sample()
    rows = c.compile_sampler(seed=3).sample(4096)
    assert rows.shape == (4096, c.num_measurements)
    varying = rows.any(axis=0) & ~rows.all(axis=0)
    # Random results vary (about half of the shots each); every other result is the reference bit.
    assert varying.sum() > 0
    np.testing.assert_array_equal(rows[:, ~varying], np.broadcast_to(ref[~varying], (4096, int((~varying).sum()))))
    frac = rows[:, varying].mean(axis=0)
    assert np.all(np.abs(frac - 0.5) < 0.1)
    # Every detector's measurements, XOR their reference bits, have even parity in every shot.
    flat = st.parse_flat(text)
    meas = 0
    for ins in flat:
        if ins.name == 'DETECTOR':
            idx = [meas - t.value for t in ins.targets if t.kind == st.REC]
            par = np.logical_xor.reduce(rows[:, idx] ^ ref[idx], axis=1)
            assert not par.any()
        meas += st.num_results(ins)


def test_skip_and_explicit_reference_sample(golden):
    import stim_amd
    c = stim_amd.Circuit(golden.circuit('surface_d3_r3'))
    ref = c.reference_sample()
    a = c.compile_sampler(seed=9).sample(10000)
    flips = c.compile_sampler(seed=9, skip_reference_sample=True).sample(10000)
    np.testing.assert_array_equal(a, flips ^ ref)
    b = c.compile_sampler(seed=9, reference_sample=ref).sample(10000)
    np.testing.assert_array_equal(a, b)
    packed = c.compile_sampler(seed=9, reference_sample=np.packbits(ref, bitorder='little')).sample(10000)
    np.testing.assert_array_equal(a, packed)
    with pytest.raises(ValueError):
        c.compile_sampler(skip_reference_sample=True, reference_sample=ref)


def test_bit_packed_and_seeds(golden):
    import stim_amd
    c = stim_amd.Circuit(golden.circuit('color_d5_r3'))
    a = c.compile_sampler(seed=4).sample(3000)
    p = c.compile_sampler(seed=4).sample(3000, bit_packed=True)
    assert p.shape == (3000, (c.num_measurements + 7) // 8) and p.dtype == np.uint8
    np.testing.assert_array_equal(_bits(p, c.num_measurements), a)
    s = c.compile_sampler(seed=4)
    first, second = s.sample(3000), s.sample(3000)
    np.testing.assert_array_equal(first, a)
    assert not np.array_equal(first, second)
    assert c.compile_sampler(seed=4).sample(0).shape == (0, c.num_measurements)


def test_sample_write_formats(golden, tmp_path):
    import stim_amd
    c = stim_amd.Circuit(golden.circuit('rep_d7_r5'))
    rows = c.compile_sampler(seed=21).sample(500)
    s = c.compile_sampler(seed=21)
    s.sample_write(500, filepath=str(tmp_path / 'm.01'), format='01')
    assert open(tmp_path / 'm.01').read() == ''.join(''.join('1' if b else '0' for b in r) + '\n' for r in rows)
    s = c.compile_sampler(seed=21)
    s.sample_write(500, filepath=str(tmp_path / 'm.dets'), format='dets')
    lines = open(tmp_path / 'm.dets').read().split('\n')[:-1]
    assert lines == ['shot' + ''.join(' M%d' % k for k in np.flatnonzero(r)) for r in rows]
    s = c.compile_sampler(seed=21)
    s.sample_write(500, filepath=str(tmp_path / 'm.b8'), format='b8')
    assert open(tmp_path / 'm.b8', 'rb').read() == np.packbits(rows, axis=1, bitorder='little').tobytes()


def test_d25_full_size_noiseless_parities(golden):
    """The d=25, 25-round circuit without noise at 2^15 shots: every detector holds in every shot."""
    import stim_amd
    from oracle import stim_text as st
    text = _gen('surface_code:rotated_memory_z', 25, 25, 0.0)
    c = stim_amd.Circuit(text)
    ref = c.reference_sample()
    m = c.num_measurements
    assert m == 16225
    packed = c.compile_sampler(seed=1).sample(1 << 15, bit_packed=True)
    rows = _bits(packed, m)
    flat = st.parse_flat(text)
    meas = 0
    checked = 0
    for ins in flat:
        if ins.name == 'DETECTOR':
            idx = [meas - t.value for t in ins.targets if t.kind == st.REC]
            assert not np.logical_xor.reduce(rows[:, idx] ^ ref[idx], axis=1).any()
            checked += 1
        meas += st.num_results(ins)
    assert checked == 15600
