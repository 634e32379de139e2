"""Detector-error-model sampler on the GPU (stim_amd.DetectorErrorModel(...).compile_sampler()).

Pinned by: the reference CLI's replay outputs (dem_replay.json, bit-exact, unfolded and REPEAT-folded
models); the oracle's replay of random masks (oracle/dem_oracle.py, itself pinned against those
outputs); and the oracle's exact per-detector rates, which a seeded sample must pass in a
Bonferroni binomial test. Thinned buckets (mechanisms of unequal probability sharing one geometric
stream) are checked on models whose probabilities span four decades.
"""

import os

import numpy as np
import pytest

from oracle import dem_oracle as do

pytestmark = pytest.mark.gpu

NAMES = ['surface_d3_r3', 'surface_d3_r3_folded', 'surface_d5_r5', 'surface_d5_r5_folded', 'rep_d7_r5',
         'rep_d7_r5_folded', 'color_d5_r3', 'color_d5_r3_folded']


def _dem(golden, name):
    with open(os.path.join(golden.dir, 'dems', name + '.dem')) as f:
        return f.read()


def _sampler(text, seed=1):
    import stim_amd
    return stim_amd.DetectorErrorModel(text).compile_sampler(seed=seed)


def _bits(a):
    return [''.join('1' if b else '0' for b in row) for row in a]


def _mask_array(masks, E):
    a = np.zeros((len(masks), E), dtype=np.bool_)
    for i, m in enumerate(masks):
        a[i, m] = True
    return a


def _z(counts, shots, p):
    p = np.clip(np.asarray(p, dtype=float), 1e-300, 1)
    return (np.asarray(counts, dtype=float) - shots * p) / np.sqrt(shots * p * (1 - p))


def _zcrit(n, alpha=1e-4):
    from scipy.stats import norm
    return abs(norm.ppf(alpha / (2 * n)))


def _check_rates(bits, rates, shots):
    """Bonferroni binomial test on every column whose rate is strictly inside (0, 1); exact columns
    (rate 0 or 1) must be constant."""
    rates = np.asarray(rates)
    counts = bits.sum(axis=0)
    fixed0 = rates == 0
    fixed1 = rates == 1
    assert np.all(counts[fixed0] == 0)
    assert np.all(counts[fixed1] == shots)
    live = ~(fixed0 | fixed1)
    z = _z(counts[live], shots, rates[live])
    assert np.max(np.abs(z), initial=0) < _zcrit(max(int(live.sum()), 1)), np.argsort(-np.abs(z))[:5]


# ---------------------------------------------------------------- replay (bit-exact)

def test_replay_matches_reference_vectors(golden):
    for c in golden.load('dem_replay.json'):
        s = _sampler(_dem(golden, c['dem']))
        assert s.num_errors == c['num_errors']
        errs = _mask_array(c['errors'], c['num_errors'])
        dets, obs, e = s.sample(len(errs), recorded_errors_to_replay=errs)
        assert e is None
        assert _bits(dets) == c['dets'], c['dem']
        assert _bits(obs) == c['obs'], c['dem']
        # Bit-packed recorded errors and bit-packed outputs give the same rows.
        packed_errs = np.packbits(errs, axis=1, bitorder='little')
        pd, po, _ = s.sample(len(errs), bit_packed=True, recorded_errors_to_replay=packed_errs)
        np.testing.assert_array_equal(np.unpackbits(pd, axis=1, bitorder='little')[:, :dets.shape[1]], dets)
        np.testing.assert_array_equal(np.unpackbits(po, axis=1, bitorder='little')[:, :obs.shape[1]], obs)


@pytest.mark.parametrize('name', ['surface_d5_r5', 'rep_d7_r5_folded', 'color_d5_r3'])
def test_replay_random_masks_match_oracle(golden, name):
    text = _dem(golden, name)
    m = do.parse(text)
    E = len(m[2])
    rng = np.random.default_rng(7)
    masks = [sorted(rng.choice(E, size=int(rng.integers(0, 13)), replace=False).tolist()) for _ in range(3000)]
    s = _sampler(text)
    dets, obs, _ = s.sample(len(masks), recorded_errors_to_replay=_mask_array(masks, E))
    want_d, want_o = do.replay(m, masks)
    np.testing.assert_array_equal(dets, want_d)
    np.testing.assert_array_equal(obs, want_o)


def _synthetic_dem(rng, n_err, D, O, pmin, pmax, max_dets=3, extra_lines=()):
    lines = list(extra_lines)
    for _ in range(n_err):
        p = float(10 ** rng.uniform(np.log10(pmin), np.log10(pmax)))
        k = int(rng.integers(0, max_dets + 1))
        ts = ['D%d' % d for d in rng.choice(D, size=k, replace=False)]
        ts += ['L%d' % o for o in range(O) if rng.random() < 0.15]
        if not ts:
            ts = ['D%d' % int(rng.integers(0, D))]
        lines.append('error(%.6g) %s' % (p, ' '.join(ts)))
    lines.append('detector D%d' % (D - 1))
    lines.append('logical_observable L%d' % (O - 1))
    return '\n'.join(lines) + '\n'


@pytest.mark.parametrize('D,O,max_dets', [(300, 2, 2), (300, 5, 4), (20000, 3, 3)])
def test_replay_synthetic_models_match_oracle(D, O, max_dets):
    """Compact records (D <= 2^14, O <= 2), their overflow entries (> 2 detectors), and wide records
    (D > 2^14 or O > 2)."""
    rng = np.random.default_rng(D + O)
    text = _synthetic_dem(rng, 4000, D, O, 1e-4, 0.3, max_dets=max_dets)
    m = do.parse(text)
    E = len(m[2])
    masks = [sorted(rng.choice(E, size=int(rng.integers(0, 20)), replace=False).tolist()) for _ in range(2500)]
    s = _sampler(text)
    dets, obs, _ = s.sample(len(masks), recorded_errors_to_replay=_mask_array(masks, E))
    want_d, want_o = do.replay(m, masks)
    np.testing.assert_array_equal(dets, want_d)
    np.testing.assert_array_equal(obs, want_o)


def test_replay_rejects_bad_shapes():
    s = _sampler('error(0.1) D0 D1\nerror(0.2) D1 L0\n')
    with pytest.raises(ValueError):
        s.sample(3, recorded_errors_to_replay=np.zeros((3, 3), dtype=np.bool_))
    with pytest.raises(ValueError):
        s.sample(3, recorded_errors_to_replay=np.zeros((2, 2), dtype=np.bool_))


# ---------------------------------------------------------------- seeded statistics

@pytest.mark.parametrize('name', ['surface_d5_r5', 'rep_d7_r5_folded', 'color_d5_r3_folded'])
def test_seeded_rates_match_exact_rates(golden, name):
    text = _dem(golden, name)
    m = do.parse(text)
    s = _sampler(text, seed=2026)
    shots = 1 << 21
    d, o, _ = s.sample(shots, bit_packed=True)
    bits = np.concatenate([np.unpackbits(d, axis=1, bitorder='little')[:, :m[0]],
                           np.unpackbits(o, axis=1, bitorder='little')[:, :m[1]]], axis=1)
    _check_rates(bits, do.exact_rates(m), shots)


def test_thinned_buckets_span_four_decades():
    """Probabilities spread over [1e-4, 0.4]: every bucket (pmax/2, pmax] mixes unequal rates, plus
    mechanisms above 1/2 and the certain and impossible ones."""
    rng = np.random.default_rng(11)
    extra = ['error(0.75) D0 D1', 'error(0.95) D2', 'error(1) D3 L0', 'error(0) D4', 'error(0.5) D5',
             'error(0.6) D6 D6 D7']
    text = _synthetic_dem(rng, 6000, 400, 2, 1e-4, 0.4, max_dets=2, extra_lines=extra)
    m = do.parse(text)
    s = _sampler(text, seed=99)
    shots = 1 << 20
    d, o, _ = s.sample(shots)
    bits = np.concatenate([d, o], axis=1)
    _check_rates(bits, do.exact_rates(m), shots)


def test_mechanism_targets_flip_together():
    s = _sampler('error(0.3) D0 D1\nerror(0.2) D1 D2 L0\nerror(0.1) D3 D3\n', seed=5)
    shots = 1 << 18
    d, o, _ = s.sample(shots)
    a, b = d[:, 0], d[:, 2]
    np.testing.assert_array_equal(d[:, 1], a ^ b)
    np.testing.assert_array_equal(o[:, 0], b)
    assert not d[:, 3].any()
    assert _z(a.sum(), shots, 0.3) ** 2 < 25
    assert _z(b.sum(), shots, 0.2) ** 2 < 25
    # Independent mechanisms: joint rate 0.06.
    assert _z((a & b).sum(), shots, 0.06) ** 2 < 25


def test_deterministic_mechanisms():
    s = _sampler('error(1) D0 L1\nerror(0) D1\nerror(1) D2 D3\nerror(1) D3\ndetector D5\n', seed=5)
    d, o, _ = s.sample(5000)
    assert d.shape == (5000, 6) and o.shape == (5000, 2)
    assert d[:, 0].all() and d[:, 2].all()
    assert not d[:, [1, 3, 4, 5]].any()
    assert o[:, 1].all() and not o[:, 0].any()


def test_seed_reproducible_and_calls_differ(golden):
    text = _dem(golden, 'surface_d5_r5')
    a = _sampler(text, seed=42)
    b = _sampler(text, seed=42)
    x1, y1, _ = a.sample(100000, bit_packed=True)
    x2, y2, _ = b.sample(100000, bit_packed=True)
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(y1, y2)
    x3, _, _ = a.sample(100000, bit_packed=True)
    assert not np.array_equal(x1, x3)
    # Unpacked rows of a fresh sampler with the same seed are the packed rows expanded.
    c = _sampler(text, seed=42)
    u, v, _ = c.sample(100000)
    np.testing.assert_array_equal(np.packbits(u, axis=1, bitorder='little'), x1)
    np.testing.assert_array_equal(np.packbits(v, axis=1, bitorder='little'), y1)


def test_shapes_and_zero_shots(golden):
    s = _sampler(_dem(golden, 'color_d5_r3'))
    d, o, e = s.sample(0)
    assert d.shape == (0, 27) and o.shape == (0, 1) and e is None
    d, o, _ = s.sample(13, bit_packed=True)
    assert d.shape == (13, 4) and o.shape == (13, 1) and d.dtype == np.uint8
    assert not (d[:, 3] >> 3).any() and not (o[:, 0] >> 1).any()   # padding bits stay zero


def test_sample_write_matches_sample(golden, tmp_path):
    text = _dem(golden, 'surface_d3_r3')
    shots = 777
    d, o, _ = _sampler(text, seed=8).sample(shots)
    s = _sampler(text, seed=8)
    s.sample_write(shots, det_out_file=str(tmp_path / 'd.01'), det_out_format='01',
                   obs_out_file=str(tmp_path / 'o.b8'), obs_out_format='b8')
    assert open(tmp_path / 'd.01').read() == ''.join(r + '\n' for r in _bits(d))
    assert open(tmp_path / 'o.b8', 'rb').read() == np.packbits(o, axis=1, bitorder='little').tobytes()
    # Replay from a file written in the reference's b8 format.
    c = golden.load('dem_replay.json')[0]
    errs = np.packbits(_mask_array(c['errors'], c['num_errors']), axis=1, bitorder='little')
    (tmp_path / 'e.b8').write_bytes(errs.tobytes())
    s.sample_write(len(errs), det_out_file=str(tmp_path / 'rd.01'), obs_out_file=str(tmp_path / 'ro.01'),
                   replay_err_in_file=str(tmp_path / 'e.b8'), replay_err_in_format='b8')
    assert open(tmp_path / 'rd.01').read().split() == c['dets']
    assert open(tmp_path / 'ro.01').read().split() == c['obs']


def _d25_text(golden):
    import gzip
    with gzip.open(os.path.join(golden.dir, 'dems', 'surface_d25_r25_folded.dem.gz'), 'rt') as f:
        return f.read()


def test_d25_model_full_size_rates(golden):
    """The d=25, 25-round model with folded loops (356,321 mechanisms once unrolled, 15,600 detectors)
    at 2^19 shots: per-detector rates against the exact rates, counted from the packed rows a bit plane
    at a time."""
    text = _d25_text(golden)
    m = do.parse(text)
    s = _sampler(text, seed=31)
    assert (s.num_detectors, s.num_observables, s.num_errors) == (15600, 1, 356321)
    n = m[0] + m[1]
    counts = np.zeros(((n + 7) // 8) * 8, dtype=np.int64)
    shots = 0
    for _ in range(2):
        d, o, _ = s.sample(1 << 18, bit_packed=True)
        for b in range(8):
            counts[b:m[0]:8] += ((d >> b) & 1).sum(axis=0, dtype=np.int64)[: len(range(b, m[0], 8))]
        counts[m[0]] += int((o[:, 0] & 1).sum())
        shots += len(d)
    assert not (d[:, -1] >> (m[0] % 8 or 8)).any()   # padding bits of the last byte stay zero
    _check_rates_counts(counts[:n], do.exact_rates(m), shots)


def _check_rates_counts(counts, rates, shots):
    z = _z(counts, shots, rates)
    assert np.max(np.abs(z)) < _zcrit(len(z)), np.argsort(-np.abs(z))[:5]


# ---------------------------------------------------------------- returned errors

def test_reference_docstring_example_returns_errors():
    """The reference's own example (dem_sampler.pybind.cc:88-110): error(0) D0, error(1) D1 D2 L0."""
    s = _sampler('error(0) D0\nerror(1) D1 D2 L0\n', seed=3)
    d, o, e = s.sample(shots=4, return_errors=True)
    np.testing.assert_array_equal(d, [[False, True, True]] * 4)
    np.testing.assert_array_equal(o, [[True]] * 4)
    np.testing.assert_array_equal(e, [[False, True]] * 4)
    d, o, e = s.sample(shots=4, return_errors=True, bit_packed=True)
    np.testing.assert_array_equal(e, [[2]] * 4)


@pytest.mark.parametrize('name', ['surface_d5_r5', 'rep_d7_r5_folded', 'color_d5_r3_folded'])
def test_returned_errors_replay_to_the_same_events(golden, name):
    """Returned errors are exactly the mechanisms that fired: replaying them (oracle and GPU replay)
    gives the same detection events and observables; the draws are those of a plain call."""
    text = _dem(golden, name)
    m = do.parse(text)
    s = _sampler(text, seed=17)
    plain = _sampler(text, seed=17)
    shots = 20000
    d, o, e = s.sample(shots, return_errors=True)
    assert e.shape == (shots, s.num_errors) and e.dtype == np.bool_
    pd, po, _ = plain.sample(shots)
    np.testing.assert_array_equal(d, pd)
    np.testing.assert_array_equal(o, po)
    od, oo = do.replay(m, [np.nonzero(r)[0].tolist() for r in e])
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(o, oo)
    rd, ro, re = _sampler(text).sample(shots, recorded_errors_to_replay=e, return_errors=True)
    np.testing.assert_array_equal(rd, d)
    np.testing.assert_array_equal(ro, o)
    np.testing.assert_array_equal(re, e)
    # Bit-packed errors are the packed bool errors (next call of both samplers).
    d2, o2, e2 = s.sample(5000, return_errors=True, bit_packed=True)
    pd2, po2, _ = plain.sample(5000, bit_packed=True)
    np.testing.assert_array_equal(d2, pd2)
    bits = np.unpackbits(e2, axis=1, bitorder='little')[:, :s.num_errors].astype(bool)
    rd2, _, _ = _sampler(text).sample(5000, recorded_errors_to_replay=bits, bit_packed=True)
    np.testing.assert_array_equal(rd2, d2)


def test_returned_error_rates_match_probabilities():
    """Each mechanism's firing rate against its probability (thinned buckets over four decades)."""
    rng = np.random.default_rng(12)
    text = _synthetic_dem(rng, 3000, 300, 2, 1e-3, 0.4, max_dets=2,
                          extra_lines=['error(1) D0', 'error(0) D1', 'error(0.75) D2 D3'])
    m = do.parse(text)
    s = _sampler(text, seed=8)
    shots = 1 << 18
    _, _, e = s.sample(shots, return_errors=True, bit_packed=True)
    counts = np.zeros(((s.num_errors + 7) // 8) * 8, dtype=np.int64)
    for b in range(8):
        counts[b::8] = ((e >> b) & 1).sum(axis=0, dtype=np.int64)
    probs = np.array([p for p, _, _ in m[2]])
    _check_rates_counts_bounded(counts[:s.num_errors], probs, shots)


def _check_rates_counts_bounded(counts, probs, shots):
    assert np.all(counts[probs == 0] == 0) and np.all(counts[probs == 1] == shots)
    live = (probs > 0) & (probs < 1)
    z = _z(counts[live], shots, probs[live])
    assert np.max(np.abs(z)) < _zcrit(int(live.sum())), np.argsort(-np.abs(z))[:5]


def test_d25_model_returned_errors_consistent(golden):
    """The d=25 model (356,321 mechanisms): errors returned for 4,096 shots replay to the same events."""
    text = _d25_text(golden)
    s = _sampler(text, seed=23)
    d, o, e = s.sample(4096, return_errors=True, bit_packed=True)
    assert e.shape == (4096, (356321 + 7) // 8)
    fired = np.unpackbits(e, axis=1, bitorder='little')[:, :356321].sum(axis=1)
    assert 60 < fired.mean() < 160
    rd, ro, _ = _sampler(text).sample(4096, recorded_errors_to_replay=e, bit_packed=True)
    np.testing.assert_array_equal(rd, d)
    np.testing.assert_array_equal(ro, o)


@pytest.mark.parametrize('fmt', ['01', 'b8', 'dets', 'hits', 'r8'])
def test_sample_write_err_out_and_replay_files(golden, tmp_path, fmt):
    from stim_amd import formats
    text = _dem(golden, 'surface_d3_r3')
    s = _sampler(text, seed=4)
    ref = _sampler(text, seed=4)
    shots = 640
    D, O, E = s.num_detectors, s.num_observables, s.num_errors
    paths = {k: str(tmp_path / f'{k}.{fmt}') for k in ('d', 'o', 'e', 'd2', 'o2')}
    s.sample_write(shots, det_out_file=paths['d'], det_out_format=fmt, obs_out_file=paths['o'], obs_out_format=fmt,
                   err_out_file=paths['e'], err_out_format=fmt)
    d, o, e = ref.sample(shots, return_errors=True)
    np.testing.assert_array_equal(formats.read_rows(paths['d'], fmt, D, [('D', D)]), d)
    np.testing.assert_array_equal(formats.read_rows(paths['o'], fmt, O, [('L', O)]), o)
    np.testing.assert_array_equal(formats.read_rows(paths['e'], fmt, E, [('M', E)]), e)
    # Replaying the written errors reproduces the written events.
    _sampler(text).sample_write(shots, det_out_file=paths['d2'], det_out_format=fmt, obs_out_file=paths['o2'],
                                obs_out_format=fmt, replay_err_in_file=paths['e'], replay_err_in_format=fmt)
    np.testing.assert_array_equal(formats.read_rows(paths['d2'], fmt, D, [('D', D)]), d)
    np.testing.assert_array_equal(formats.read_rows(paths['o2'], fmt, O, [('L', O)]), o)
