"""Reference samples (stim_amd.Circuit.reference_sample, the host tableau behind measurement sampling
and measurement-to-detection-event conversion).

Pinned by: the reference CLI's own reference samples (ref_samples.json, written by
tests/golden/make_fixtures.py through `stim m2d` on a copy of each circuit whose detectors are its
single measurements) for the fixture circuits (up to the d=25, 25-round surface code), the gate zoo,
and 500 seeded circuits built to make signs and collapse order matter (random Clifford layers,
measurements and resets in every basis with inverted results, Pauli-product measurements and phases
with repeated factors, multi-pair MXX/MYY/MZZ, feedback, MPAD, heralded noise). Bit-exact. No GPU is used.
"""

import numpy as np
import pytest


def _text(golden, c):
    return c.get('circuit') or golden.circuit(c['name'])


def test_reference_samples_match_reference_cli(golden):
    import stim_amd
    cases = golden.load('ref_samples.json')
    assert len(cases) >= 500
    bad = []
    for c in cases:
        mine = ''.join('1' if b else '0' for b in stim_amd.Circuit(_text(golden, c)).reference_sample())
        if mine != c['ref']:
            bad.append(c['name'])
    assert not bad, bad[:10]
    # The fixtures are not trivially zero: random collapses and inverted results give ones.
    assert sum(c['ref'].count('1') for c in cases) > 2000


def test_reference_sample_bit_packed(golden):
    import stim_amd
    c = stim_amd.Circuit(golden.circuit('color_d5_r3'))
    full = c.reference_sample()
    packed = c.reference_sample(bit_packed=True)
    assert full.dtype == np.bool_ and full.shape == (c.num_measurements,)
    assert packed.dtype == np.uint8 and packed.shape == ((c.num_measurements + 7) // 8,)
    np.testing.assert_array_equal(np.unpackbits(packed, bitorder='little')[:len(full)].astype(bool), full)


def test_reference_sample_small_cases():
    import stim_amd
    assert stim_amd.Circuit('X 1\nM 0 1').reference_sample().tolist() == [False, True]
    assert stim_amd.Circuit('').reference_sample().shape == (0,)
    # Random results are 0; a correlated partner is then fixed (Bell pair measured twice over).
    assert stim_amd.Circuit('H 0\nCX 0 1\nM 0 1\nMX 0').reference_sample().tolist() == [False, False, False]
    assert stim_amd.Circuit('H 0\nCX 0 1\nX 1\nM 1 0').reference_sample().tolist() == [False, True]
    # MPAD and inverted results.
    assert stim_amd.Circuit('MPAD 1 0\nM !0').reference_sample().tolist() == [True, False, True]


@pytest.mark.parametrize('bad', ['MPP X0*Z0', 'SPP X0*rec[-1]', 'M 0\nCX 0 rec[-1]'])
def test_reference_sample_rejects_invalid_circuits(bad):
    import stim_amd
    with pytest.raises(ValueError):
        stim_amd.Circuit(bad).reference_sample()


def test_reference_sample_tableau_limit():
    """The reference sample's stabilizer tableau grows as the square of the qubit count: a circuit on
    200,001 qubits (about 20 GB of tableau) is refused with RuntimeError before allocating, and a
    circuit without measurements needs no tableau at all."""
    import stim_amd
    with pytest.raises(RuntimeError, match='tableau'):
        stim_amd.Circuit('H 200000\nM 0\n').reference_sample()
    assert len(stim_amd.Circuit('H 5000000\nX 3\n').reference_sample()) == 0
