"""Measurement sampler on the GPU (stim_amd.Circuit.compile_sampler()), on both of its paths: the
frame engine, and the expansion of detection events drawn on the symptom engine into records.

Pinned by: the reference CLI's seeded `stim sample` counts per measurement (meas_counts.json; a
two-sample binomial test, Bonferroni over every measurement), the reference sample (itself pinned
bit-exact against the reference CLI) for every deterministic result of noiseless circuits, the
parity of every detector of a noiseless circuit (its measurements XOR their reference bits is 0),
and, for the expansion, the detection events of its records against the oracle's exact rates.
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


ENGINES = ['frame', 'expansion']
# Fixtures whose detectors and observables fix every deterministic parity: the expansion applies.
EXPANDABLE = ['surface_d3_r3', 'rep_d7_r5', 'color_d5_r3', 'unrotated_d3_r2', 'surface_x_d5_r3']


def _sampler(c, engine, **kw):
    s = c.compile_sampler(**kw)
    s.set_engine(engine)
    return s


@pytest.mark.parametrize('engine', ['frame', 'auto'])
def test_counts_match_reference_cli(golden, engine):
    import stim_amd
    fixtures = golden.load('meas_counts.json')
    zs = []
    expanded = []
    for name, f in fixtures.items():
        text = f['circuit'] or golden.circuit(name)
        s = _sampler(stim_amd.Circuit(text), engine, seed=11)
        shots = 1 << 20
        m = s.num_measurements
        assert m == len(f['counts'])
        ours = _bits(s.sample(shots, bit_packed=True), m).sum(axis=0)
        if s.engine_info()['engine'] == 'expansion':
            expanded.append(name)
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
    if engine == 'auto':
        # Calls this large expand wherever the expansion applies.
        assert set(EXPANDABLE) <= set(expanded), expanded
    else:
        assert not expanded


@pytest.mark.parametrize('engine', ENGINES)
def test_noiseless_results_follow_the_reference_sample(engine):
    import stim_amd
    from oracle import stim_text as st
    text = _gen('surface_code:rotated_memory_x', 5, 4, 0.0)
    c = stim_amd.Circuit(text)
    ref = c.reference_sample()
    rows = _sampler(c, engine, seed=3).sample(4096)
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


@pytest.mark.parametrize('engine', ENGINES)
def test_bit_packed_and_seeds(golden, engine):
    import stim_amd
    c = stim_amd.Circuit(golden.circuit('color_d5_r3'))
    a = _sampler(c, engine, seed=4).sample(3000)
    p = _sampler(c, engine, seed=4).sample(3000, bit_packed=True)
    assert p.shape == (3000, (c.num_measurements + 7) // 8) and p.dtype == np.uint8
    np.testing.assert_array_equal(_bits(p, c.num_measurements), a)
    s = _sampler(c, engine, seed=4)
    first, second = s.sample(3000), s.sample(3000)
    np.testing.assert_array_equal(first, a)
    assert not np.array_equal(first, second)
    assert _sampler(c, engine, seed=4).sample(0).shape == (0, c.num_measurements)


@pytest.mark.parametrize('engine', ENGINES)
def test_sample_write_formats(golden, tmp_path, engine):
    import stim_amd
    c = stim_amd.Circuit(golden.circuit('rep_d7_r5'))
    rows = _sampler(c, engine, seed=21).sample(500)
    s = _sampler(c, engine, seed=21)
    s.sample_write(500, filepath=str(tmp_path / 'm.01'), format='01')
    assert open(tmp_path / 'm.01').read() == ''.join(''.join('1' if b else '0' for b in r) + '\n' for r in rows)
    s = _sampler(c, engine, seed=21)
    s.sample_write(500, filepath=str(tmp_path / 'm.dets'), format='dets')
    lines = open(tmp_path / 'm.dets').read().split('\n')[:-1]
    assert lines == ['shot' + ''.join(' M%d' % k for k in np.flatnonzero(r)) for r in rows]
    s = _sampler(c, engine, seed=21)
    s.sample_write(500, filepath=str(tmp_path / 'm.b8'), format='b8')
    assert open(tmp_path / 'm.b8', 'rb').read() == np.packbits(rows, axis=1, bitorder='little').tobytes()


@pytest.mark.parametrize('engine', ENGINES)
def test_d25_full_size_noiseless_parities(golden, engine):
    """The d=25, 25-round circuit without noise at 2^15 shots: every detector holds in every shot."""
    import stim_amd
    from oracle import stim_text as st
    text = _gen('surface_code:rotated_memory_z', 25, 25, 0.0)
    c = stim_amd.Circuit(text)
    ref = c.reference_sample()
    m = c.num_measurements
    assert m == 16225
    packed = _sampler(c, engine, seed=1).sample(1 << 15, bit_packed=True)
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


# ---------------------------------------------------------------- expansion path

@pytest.mark.parametrize('name', ['surface_d5_r5', 'color_d5_r3', 'surface_x_d5_r3', 'rep_d7_r5'])
def test_expansion_detection_events_match_exact_rates(golden, name):
    """Records from the expansion, converted back to detection events and observable flips by the
    oracle's m2d restatement: per-column rates against the oracle's exact rates."""
    import stim_amd
    from oracle import frame_oracle as fo
    from oracle import m2d_oracle
    text = golden.circuit(name)
    c = stim_amd.Circuit(text)
    s = _sampler(c, 'expansion', seed=5)
    shots = 1 << 18
    rows = _bits(s.sample(shots, bit_packed=True), c.num_measurements)
    y = m2d_oracle.convert(text, rows, c.reference_sample())
    exact = np.asarray(fo.exact_detector_rates(text), dtype=float)
    assert y.shape[1] == len(exact)
    p = np.clip(exact, 1e-300, 1)
    z = (y.sum(axis=0) - shots * p) / np.sqrt(shots * p * (1 - p))
    assert np.max(np.abs(z)) < _zcrit(len(z)), (name, np.argsort(-np.abs(z))[:5])


@pytest.mark.parametrize('name', ['surface_d5_r5', 'surface_x_d5_r3', 'color_d5_r3'])
def test_expansion_marginals_and_pairs_agree_with_frame_engine(golden, name):
    """Per-result rates and the rates of pairs of results (every result with the next one and with
    the one a round later) on both paths, two-sample tests: the random results' joint structure (a
    first-round random check repeats in later rounds) must match the frame engine's."""
    import stim_amd
    c = stim_amd.Circuit(golden.circuit(name))
    m = c.num_measurements
    shots = 1 << 18
    a = _bits(_sampler(c, 'frame', seed=6).sample(shots, bit_packed=True), m)
    b = _bits(_sampler(c, 'expansion', seed=7).sample(shots, bit_packed=True), m)
    stats_a, stats_b = [a.mean(axis=0)], [b.mean(axis=0)]
    for lag in (1, 8, 24):
        if lag < m:
            stats_a.append((a[:, lag:] ^ a[:, :-lag]).mean(axis=0))
            stats_b.append((b[:, lag:] ^ b[:, :-lag]).mean(axis=0))
    pa, pb = np.concatenate(stats_a), np.concatenate(stats_b)
    pool = (pa + pb) / 2
    keep = (pool > 0) & (pool < 1)
    z = (pa[keep] - pb[keep]) / np.sqrt(pool[keep] * (1 - pool[keep]) * 2 / shots)
    assert np.max(np.abs(z)) < _zcrit(len(z)), name


def test_expansion_d25_full_size_rates(golden):
    """The headline circuit at 2^18 shots on the expansion path, records left in HBM and converted to
    detection events there (the m2d kernel, itself pinned on the reference CLI's vectors): every
    detector's rate and the observable's against the exact rates of the circuit's folded error model."""
    import gzip
    import os
    import torch
    import stim_amd
    from stim_amd import _native as N
    from oracle import dem_oracle as do
    c = stim_amd.Circuit(golden.circuit('surface_d25_r25'))
    s = _sampler(c, 'expansion', seed=8)
    info = s.engine_info()
    assert info['engine'] == 'expansion' and info['free_results'] == 624, info
    shots, M = 1 << 18, c.num_measurements
    row = (M + 7) // 8
    meas = torch.empty((shots, row), dtype=torch.uint8, device='cuda')
    N.check(N.lib().mfs_sample(s._h, shots, N.MFS_BIT_PACKED | N.MFS_DEVICE_OUTPUT, meas.data_ptr(), row, None, 0))
    conv = c.compile_m2d_converter()
    D, O = conv.num_detectors, conv.num_observables
    out = torch.empty((shots, (D + O + 7) // 8), dtype=torch.uint8, device='cuda')
    N.check(N.lib().mfs_m2d_convert(conv._h, shots, meas.data_ptr(), row,
                                    N.MFS_BIT_PACKED | N.MFS_APPEND_OBSERVABLES | N.MFS_DEVICE_OUTPUT,
                                    out.data_ptr(), out.shape[1], None, 0))
    counts = torch.zeros(out.shape[1] * 8, dtype=torch.int64, device='cuda')
    for b in range(8):
        counts[b::8] = ((out >> b) & 1).sum(dim=0, dtype=torch.int64)
    counts = counts[:D + O].cpu().numpy()
    del meas, out
    with gzip.open(os.path.join(golden.dir, 'dems', 'surface_d25_r25_folded.dem.gz'), 'rt') as f:
        exact = np.asarray(do.exact_rates(do.parse(f.read())), dtype=float)
    p = np.clip(exact, 1e-300, 1)
    z = (counts - shots * p) / np.sqrt(shots * p * (1 - p))
    assert np.max(np.abs(z)) < _zcrit(len(z)), np.argsort(-np.abs(z))[:5]
    torch.cuda.synchronize()
