"""The measurement sampler's expansion plan, on the host (no GPU).

Where a circuit's detectors and observables fix every parity its noiseless run fixes, a measurement
record is determined by its detection events and observable flips up to the random results. The
planner checks that by dimension (unconstrained results = random results of the reference run); the
host restatement of the expansion must then map the detection events
of ANY record x back to a record m with the same detection events and observable flips, whatever
values the unconstrained results take. Detection events are computed by the oracle's restatement of
the reference's measurements-to-detection-events conversion, which the reference CLI's `stim m2d`
vectors pin (test_m2d_oracle.py). Circuits the plan must refuse are refused.
"""

import ctypes

import numpy as np
import pytest

from oracle import m2d_oracle


def _lib():
    from stim_amd import _native
    return _native, _native.lib()


def _stats(text):
    N, L = _lib()
    out = (ctypes.c_uint64 * 8)()
    b = text.encode()
    N.check(L.mfs_expand_plan_stats(b, len(b), out))
    return dict(zip(['complete', 'M', 'pivots', 'free', 'random', 'random_total', 'phases', 'depth'], list(out)))


def _expand(text, y, free):
    N, L = _lib()
    y = np.ascontiguousarray(y, dtype=np.uint8)
    free = np.ascontiguousarray(free, dtype=np.uint8)
    M = _stats(text)['M']
    out = np.zeros((y.shape[0], M), dtype=np.uint8)
    b = text.encode()
    N.check(L.mfs_expand_shots_host(b, len(b), y.shape[0], y.ctypes.data, max(y.shape[1], 1),
                                    free.ctypes.data if free.size else None, max(free.shape[1], 1),
                                    out.ctypes.data, max(M, 1)))
    return out.astype(bool)


MEMORY = ['surface_d3_r3', 'surface_d5_r5', 'rep_d7_r5', 'surface_x_d5_r3', 'color_d5_r3']


@pytest.mark.parametrize('name', MEMORY + ['surface_d25_r25'])
def test_plan_counts(golden, name):
    s = _stats(golden.circuit(name))
    assert s['complete'] == 1, s
    assert s['free'] == s['random'] == s['M'] - s['pivots']


def test_plan_applies_to_the_headline_circuit(golden):
    """The d=25, 25-round memory-Z circuit: 624 random results (312 first-round X checks, 312
    degrees of freedom of the final data measurements), 15,601 independent rows."""
    s = _stats(golden.circuit('surface_d25_r25'))
    assert s['complete'] == 1, s
    assert s['M'] == 16225 and s['random'] == 624 and s['pivots'] == 16225 - 624
    # One chain per check through the 25 rounds, and runs along the data qubits whose other terms are
    # XORed in by a phase of independent records before each: five barrier phases for a dependency
    # depth of 48.
    assert s['phases'] == 5 and s['depth'] == 48, s


@pytest.mark.parametrize('name', MEMORY)
def test_expansion_reproduces_detection_events(golden, name):
    text = golden.circuit(name)
    s = _stats(text)
    if not s['complete']:
        pytest.skip(f'expansion does not apply to {name}')
    rng = np.random.default_rng(11)
    shots, M = 48, s['M']
    x = rng.integers(0, 2, size=(shots, M)).astype(bool)
    zero = np.zeros(M, dtype=bool)
    y = m2d_oracle.convert(text, x, zero)
    for trial in range(2):
        free = rng.integers(0, 2, size=(shots, s['free'])).astype(np.uint8)
        m = _expand(text, y, free)
        np.testing.assert_array_equal(m2d_oracle.convert(text, m, zero), y, err_msg=f'{name} trial {trial}')
    # With the unconstrained results fixed to x's own values the expansion returns x itself.
    free_idx = _free_indices(text, M)
    m = _expand(text, y, x[:, free_idx].astype(np.uint8))
    np.testing.assert_array_equal(m, x)


GENERATED = [
    ('surface_code:unrotated_memory_z', 5, 4),
    ('surface_code:unrotated_memory_x', 5, 4),
    ('surface_code:rotated_memory_x', 7, 7),
    ('surface_code:rotated_memory_z', 9, 3),
    ('surface_code:rotated_memory_z', 11, 11),
    ('repetition_code:memory', 11, 9),
]


@pytest.mark.parametrize('code,d,r', GENERATED)
def test_expansion_of_generated_circuits(code, d, r):
    """Generated memory experiments whose plans have substitution runs with pre-phases (five phases for
    the surface codes): any record x is reproduced up to its free results."""
    from stim_amd import gen
    text = gen.generate(code, distance=d, rounds=r, after_clifford_depolarization=1e-3,
                        before_measure_flip_probability=1e-3)
    s = _stats(text)
    assert s['complete'] == 1 and s['free'] == s['random'], s
    if code.startswith('surface_code') and d >= 7:
        assert s['phases'] == 5, s
    rng = np.random.default_rng(d * 100 + r)
    M = s['M']
    x = rng.integers(0, 2, size=(32, M)).astype(bool)
    zero = np.zeros(M, dtype=bool)
    y = m2d_oracle.convert(text, x, zero)
    m = _expand(text, y, rng.integers(0, 2, size=(32, s['free'])).astype(np.uint8))
    np.testing.assert_array_equal(m2d_oracle.convert(text, m, zero), y)


def _free_indices(text, M):
    """Unconstrained results: the record indices that no row pivots on, found by expanding the zero
    record with one free bit set at a time."""
    F = _stats(text)['free']
    _, dets, obs = m2d_oracle.columns(text)
    y = np.zeros((F, len(dets) + len(obs)), dtype=np.uint8)
    m = _expand(text, y, np.eye(F, dtype=np.uint8))
    idx = []
    for f in range(F):
        ones = np.flatnonzero(m[f])
        idx.append(int(ones[0]))   # a free result's own slot is its first (lowest) effect
    return np.array(idx, dtype=np.int64)


@pytest.mark.parametrize('text,reason', [
    ('H 0\nCX 0 1\nR 0\nM 1\n', 'a random reset leaks into a later result'),
    ('X_ERROR(0.1) 0\nM 0\n', 'a noisy result no detector covers'),
    ('R 0 1\nX_ERROR(0.1) 0\nM 0 1\nDETECTOR rec[-2]\n', 'result 1 is deterministic but undetected'),
])
def test_plan_refuses(text, reason):
    assert _stats(text)['complete'] == 0, reason


def test_plan_accepts_all_random_and_fully_detected_records():
    # Every result random (no constraint), and a deterministic record whose results are all detected.
    assert _stats('H 0 1\nM 0 1\n')['complete'] == 1
    # A random reset of a qubit that shares no state with the rest leaks nothing.
    assert _stats('RX 0\nM 0\n')['complete'] == 1
    s = _stats('R 0 1\nX_ERROR(0.1) 0 1\nM 0 1\nDETECTOR rec[-1]\nDETECTOR rec[-2] rec[-1]\n')
    assert s['complete'] == 1 and s['free'] == 0
