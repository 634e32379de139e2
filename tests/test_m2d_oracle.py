"""Pins the conversion oracle (oracle/m2d_oracle.py) against the reference CLI's `stim m2d` outputs
(m2d_vectors.json: seeded random measurement data for the fixture circuits up to d=25, the gate zoo,
and sign-stress circuits with random detectors and observables; with and without the reference
sample). A planted mistake in the oracle (dropping the reference parity) must be caught. No GPU.
"""

import numpy as np
import pytest

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


# measurements_to_detection_events.test.cc:25-143 (single_detector_no_sweep_data): shot-0 measurement
# bits, circuit, expected detector 0. Conversion against the circuit's own reference sample (the host
# tableau, no GPU needed) checks the oracle and the reference sample together.
SINGLE_DETECTOR = [
    ((0,), 'M 0\nDETECTOR rec[-1]\n', 0),
    ((0,), 'M !0\nDETECTOR rec[-1]\n', 1),
    ((0,), 'X 0\nM 0\nDETECTOR rec[-1]\n', 1),
    ((1,), 'M 0\nDETECTOR rec[-1]\n', 1),
    ((1,), 'M !0\nDETECTOR rec[-1]\n', 0),
    ((1,), 'X 0\nM 0\nDETECTOR rec[-1]\n', 0),
    ((1, 0), 'M 0 1\nDETECTOR rec[-1]\n', 0),
    ((1, 0), 'M 0 1\nDETECTOR rec[-2]\n', 1),
    ((1, 0), 'M 0 1\nDETECTOR rec[-1] rec[-2]\n', 1),
    ((1, 0), 'M 0 1\nDETECTOR\n', 0),
]


@pytest.mark.parametrize('meas,text,want', SINGLE_DETECTOR)
def test_reference_single_detector_answers(meas, text, want):
    import stim_amd
    ref = np.asarray(stim_amd.Circuit(text).reference_sample(), dtype=bool)
    out = mo.convert(text, np.array([meas], dtype=bool), ref)
    assert out.shape == (1, 1) and bool(out[0, 0]) == bool(want)


def test_reference_appended_observable_answers():
    """measurements_to_detection_events.test.cc:299-371 (append_observables): all-zero measurements;
    observable 9 lands at column D + 9 when appended."""
    import stim_amd
    text = 'X 0\nM 0\nOBSERVABLE_INCLUDE(9) rec[-1]\n'
    ref = np.asarray(stim_amd.Circuit(text).reference_sample(), dtype=bool)
    out = mo.convert(text, np.zeros((1, 1), dtype=bool), ref)
    assert out.shape == (1, 10) and out[0, 9] and not out[0, :9].any()
    text = 'X 0\nM 0\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(9) rec[-1]\nDETECTOR rec[-1]\n'
    ref = np.asarray(stim_amd.Circuit(text).reference_sample(), dtype=bool)
    out = mo.convert(text, np.zeros((1, 1), dtype=bool), ref)
    assert out.shape == (1, 12)
    assert out[0, 0] and out[0, 1] and not out[0, 2:11].any() and out[0, 11]
