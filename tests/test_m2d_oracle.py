"""Pins the conversion oracle (oracle/m2d_oracle.py) against the reference CLI's `stim m2d` outputs
(m2d_vectors.json: seeded random measurement data for the fixture circuits up to d=25, the gate zoo,
and sign-stress circuits with random detectors and observables; with and without the reference
sample). A planted mistake in the oracle (dropping the reference parity) must be caught. No GPU.
"""

import numpy as np

from oracle import m2d_oracle as mo


def _rows(hexdata, shots, bits):
    raw = np.frombuffer(bytes.fromhex(hexdata), dtype=np.uint8).reshape(shots, -1) if shots else np.zeros((0, 0), np.uint8)
    return np.unpackbits(raw, axis=1, bitorder='little')[:, :bits].astype(bool)


def _cases(golden):
    for c in golden.load('m2d_vectors.json'):
        text = c['circuit'] or golden.circuit(c['name'])
        m, dets, obs = mo.columns(text)
        yield c, text, m, len(dets) + len(obs)


def test_oracle_matches_reference_conversions(golden):
    n = 0
    for c, text, m, ncols in _cases(golden):
        meas = _rows(c['measurements'], c['shots'], m)
        ref = np.zeros(m, bool) if c['skip_reference_sample'] else np.array([b == '1' for b in c['reference_sample']], bool)
        want = _rows(c['expected'], c['shots'], ncols)
        np.testing.assert_array_equal(mo.convert(text, meas, ref), want, err_msg=c['name'])
        n += 1
    assert n >= 90


def test_planted_mistake_is_caught(golden):
    caught = 0
    for c, text, m, ncols in _cases(golden):
        if c['skip_reference_sample'] or '1' not in c['reference_sample']:
            continue
        meas = _rows(c['measurements'], c['shots'], m)
        want = _rows(c['expected'], c['shots'], ncols)
        caught += not np.array_equal(mo.convert(text, meas, np.zeros(m, bool)), want)
    assert caught >= 10
