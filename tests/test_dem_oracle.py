"""Pins the detector-error-model oracle (oracle/dem_oracle.py) and checks the library's DEM parser.

Checks: the reference CLI's replay outputs (`sample_dem --replay_err_in`) for unfolded and
REPEAT-folded models; exact rates of every model against the circuit-level rates; folded and
unfolded models of one circuit give the same rates; the rate formula against brute-force
enumeration; the library's parse counts against the oracle's; the library rejects malformed text.
No GPU is used.
"""

import ctypes
import itertools
import math
import os

import numpy as np
import pytest

from oracle import dem_oracle as do


def _dem(golden, name):
    with open(os.path.join(golden.dir, 'dems', name + '.dem')) as f:
        return f.read()


def _bits(a):
    return [''.join('1' if b else '0' for b in row) for row in a]


def test_oracle_replay_matches_reference(golden):
    cases = golden.load('dem_replay.json')
    assert len(cases) == 8
    assert any('repeat' in _dem(golden, c['dem']) for c in cases)
    for c in cases:
        m = do.parse(_dem(golden, c['dem']))
        assert len(m[2]) == c['num_errors'], c['dem']
        dets, obs = do.replay(m, c['errors'])
        assert _bits(dets) == c['dets'], c['dem']
        assert _bits(obs) == c['obs'], c['dem']


def test_oracle_rates_match_circuit_rates(golden):
    rates = golden.load('dem_rates.json')
    for name in ['surface_d3_r3', 'surface_d5_r5', 'rep_d7_r5', 'color_d5_r3']:
        for suffix in ('', '_folded'):
            m = do.parse(_dem(golden, name + suffix))
            np.testing.assert_allclose(do.exact_rates(m), rates[name], rtol=1e-9, atol=1e-15,
                                       err_msg=name + suffix)


def test_folded_and_unfolded_models_agree(golden):
    for name in ['surface_d3_r3', 'surface_d5_r5', 'rep_d7_r5', 'color_d5_r3']:
        a = do.parse(_dem(golden, name))
        b = do.parse(_dem(golden, name + '_folded'))
        assert a[:2] == b[:2]
        np.testing.assert_allclose(do.exact_rates(a), do.exact_rates(b), rtol=1e-12)


def _brute_rates(model):
    nd, no, errors = model
    out = np.zeros(nd + no)
    for fired in itertools.product([0, 1], repeat=len(errors)):
        w = 1.0
        for f, (p, _, _) in zip(fired, errors):
            w *= p if f else 1 - p
        d, o = do.replay(model, [[e for e, f in enumerate(fired) if f]])
        out += w * np.concatenate([d[0], o[0]])
    return out


@pytest.mark.parametrize('text', [
    'error(0.1) D0 D1\nerror(0.2) D1 L0\nerror(0.7) D0 D2\nerror(0.5) D2',
    'error(0.3) D0 D0 D1\nerror(1) D1 L1\nerror(0) D0\nshift_detectors 2\nerror(0.25) D0 L0',
    'repeat 3 {\n  error(0.05) D0 D1\n  shift_detectors 1\n}\nerror(0.9) D0 L0 L0',
])
def test_exact_rates_match_brute_force(text):
    m = do.parse(text)
    np.testing.assert_allclose(do.exact_rates(m), _brute_rates(m), rtol=1e-12, atol=1e-15)


def test_parser_semantics_small():
    nd, no, errors = do.parse('error(0.3) D0 D0 D1 ^ D4 L2\ndetector(1, 2) D9\nshift_detectors(1.5) 3\n'
                              'error(0.1) D1 # comment\nlogical_observable L5\nerror[tag](0.2) L1 L1')
    assert (nd, no) == (10, 6)
    assert errors == [(0.3, [1, 4], [2]), (0.1, [4], []), (0.2, [], [])]


def _lib():
    from stim_amd import _native as N
    return N, N.lib()


def _stats(text):
    N, lib = _lib()
    out = (ctypes.c_uint64 * 3)()
    t = text.encode()
    rc = lib.mfs_dem_parse_stats(t, len(t), out)
    return rc, list(out)


def test_library_parse_counts_match_oracle(golden):
    names = sorted(f[:-4] for f in os.listdir(os.path.join(golden.dir, 'dems')) if f.endswith('.dem'))
    assert len(names) == 8
    for name in names:
        text = _dem(golden, name)
        nd, no, errors = do.parse(text)
        rc, out = _stats(text)
        assert rc == 0
        assert out == [nd, no, len(errors)], name
    rc, out = _stats('error(0.3) D0 D0 D1 ^ D4 L2\ndetector(1, 2) D9\nshift_detectors(1.5) 3\n'
                     'error(0.1) D1 # comment\nlogical_observable L5\nerror[tag](0.2) L1 L1')
    assert rc == 0 and out == [10, 6, 3]


@pytest.mark.parametrize('bad', [
    'error(1.5) D0', 'error(-0.1) D0', 'error D0', 'error(0.1, 0.2) D0', 'error(0.1) 5',
    'detector L0', 'logical_observable D0', 'repeat 2 {\nerror(0.1) D0', '}', 'foo D0',
    'repeat {\n}', 'error(0.1) Dx', 'error(0.1)D0',
])
def test_library_rejects_malformed_models(bad):
    N, lib = _lib()
    rc, _ = _stats(bad)
    assert rc == N.MFS_ERR_INVALID_ARGUMENT, bad
    assert b'detector error model line' in lib.mfs_last_error()


def test_library_parse_counts_match_oracle_d25(golden):
    import gzip
    with gzip.open(os.path.join(golden.dir, 'dems', 'surface_d25_r25_folded.dem.gz'), 'rt') as f:
        text = f.read()
    assert 'repeat' in text
    nd, no, errors = do.parse(text)
    assert (nd, no, len(errors)) == (15600, 1, 356321)
    rc, out = _stats(text)
    assert rc == 0 and out == [nd, no, len(errors)]


def test_reference_basic_probability_model_rates():
    """dem_sampler.test.cc:50-75 (resample_basic_probabilities): the model's detectors fire at exactly
    0, 1/4, 1/2, 3/4, 1, 1 (the '^' separator still flips both D4 and D5) and L0 with D1."""
    from oracle import dem_oracle as do
    model = do.parse('error(0) D0\nerror(0.25) D1 L0\nerror(0.5) D2\nerror(0.75) D3\nerror(1) D4 ^ D5\n')
    nd, no, _ = model
    assert (nd, no) == (6, 1)
    np.testing.assert_allclose(do.exact_rates(model), [0, 0.25, 0.5, 0.75, 1, 1, 0.25], atol=1e-12)
    # Replaying mechanism 1 flips D1 and L0; mechanism 4 flips D4 and D5.
    dets, obs = do.replay(model, [[1], [4], [0, 1, 2, 3, 4]])
    assert dets.tolist() == [[False, True, False, False, False, False],
                             [False, False, False, False, True, True],
                             [True] * 6]
    assert obs.tolist() == [[True], [False], [True]]
