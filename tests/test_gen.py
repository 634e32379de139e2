"""The workload generators: the reference generator's circuit text, reproduced exactly.

Pins: tests/golden/gen_pins.json holds `stim gen` output (reference gen/gen_surface_code.cc,
gen/gen_rep_code.cc) for 14 parameter sets, including one round, two rounds, zero noise, unequal noise
strengths and arguments that print with six significant digits."""

import json
import os

import pytest

from stim_amd import gen
import stim_amd

HERE = os.path.dirname(os.path.abspath(__file__))


def _strip_comments(text):
    return '\n'.join(l for l in text.splitlines() if not l.startswith('#')) + '\n'


def _pins():
    return json.load(open(os.path.join(HERE, 'golden', 'gen_pins.json')))


@pytest.mark.parametrize('pin', _pins(), ids=lambda e: f"{e['code']}-{e['task']}-d{e['distance']}-r{e['rounds']}")
def test_generator_text_matches_reference(pin):
    got = gen.generate(f"{pin['code']}:{pin['task']}", distance=pin['distance'], rounds=pin['rounds'],
                       **pin['params'])
    assert got == _strip_comments(pin['text'])


def test_generated_circuit_counts():
    # Rotated memory: D = (d^2 - 1) r, M = (d^2 - 1) r + d^2; repetition: D = (d - 1)(r + 1).
    for d, r in [(3, 3), (5, 2), (9, 9)]:
        c = stim_amd.Circuit.generated('surface_code:rotated_memory_z', distance=d, rounds=r,
                                       after_clifford_depolarization=1e-3)
        assert c.num_detectors == (d * d - 1) * r
        assert c.num_measurements == (d * d - 1) * r + d * d
        assert c.num_observables == 1
    c = stim_amd.Circuit.generated('repetition_code:memory', distance=7, rounds=4)
    assert c.num_detectors == 6 * 5


def test_generator_argument_errors():
    with pytest.raises(ValueError):
        gen.generate('surface_code:rotated_memory_z', distance=1, rounds=3)
    with pytest.raises(ValueError):
        gen.generate('surface_code:rotated_memory_z', distance=3, rounds=0)
    with pytest.raises(ValueError):
        gen.generate('surface_code:rotated_memory_y', distance=3, rounds=3)
    with pytest.raises(ValueError):
        gen.generate('surface_code:rotated_memory_z', distance=3, rounds=3, after_clifford_depolarization=1.5)


def test_substitute_probability_round_trip():
    base = gen.generate('surface_code:rotated_memory_z', distance=5, rounds=5, after_clifford_depolarization=1e-3,
                        after_reset_flip_probability=1e-3, before_measure_flip_probability=1e-3)
    want = gen.generate('surface_code:rotated_memory_z', distance=5, rounds=5, after_clifford_depolarization=2e-3,
                        after_reset_flip_probability=2e-3, before_measure_flip_probability=2e-3)
    assert gen.substitute_probability(base, 0.001, 0.002) == want
