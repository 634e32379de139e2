"""Pins the oracle (oracle/frame_oracle.py) before anything is checked against it.

Checks: every fixed-fault vector the reference CLI produced; per-detector rates implied by the
reference's detector error model; the reference's own seeded firing counts; and a planted mistake
in the oracle's CX rule must be caught by the fixtures (so they can catch such mistakes at all).
"""

import math

import numpy as np
import pytest

from oracle import frame_oracle as fo
from oracle import stim_text as st


def _bits(dets, obs, i):
    return ''.join('1' if b else '0' for b in np.concatenate([dets[i], obs[i]]))


def _mismatches(fixtures, oracle_cls=fo.FrameOracle):
    bad = 0
    total = 0
    for f in fixtures:
        o = oracle_cls(f['circuit'])
        sets = f['sets']
        dets, obs = o.run(len(sets), faults=[[tuple(x) for x in s['faults']] for s in sets])
        for i, s in enumerate(sets):
            total += 1
            bad += _bits(dets, obs, i) != s['expected']
    return bad, total


def test_oracle_matches_reference_fault_vectors(golden):
    fixtures = golden.load('fault_vectors.json')
    bad, total = _mismatches(fixtures)
    assert total >= 600
    assert bad == 0


def test_oracle_stats_match_fixture_stats(golden):
    for f in golden.load('fault_vectors.json'):
        s = st.stats(st.parse_flat(f['circuit']))
        assert s == f['stats']


def test_exact_rates_match_reference_dem(golden):
    rates = golden.load('dem_rates.json')
    for name, r in rates.items():
        mine = fo.exact_detector_rates(golden.circuit(name))
        np.testing.assert_allclose(mine, r, rtol=1e-9, atol=1e-15)


def _binomial_z(counts, shots, p):
    p = np.clip(np.asarray(p, dtype=float), 1e-300, 1)
    mu = shots * p
    sd = np.sqrt(shots * p * (1 - p))
    return (np.asarray(counts) - mu) / sd


def test_reference_seeded_counts_agree_with_exact_rates(golden):
    """The reference sampler's own firing counts pass the same binomial test the GPU must pass."""
    counts = golden.load('seeded_counts.json')
    for name, c in counts.items():
        exact = fo.exact_detector_rates(golden.circuit(name))
        z = _binomial_z(c['counts'], c['shots'], exact)
        # Bonferroni over all detectors at family-wise alpha = 1e-4.
        zcrit = abs(_norm_ppf(1e-4 / (2 * len(z))))
        assert np.max(np.abs(z)) < zcrit, (name, np.max(np.abs(z)))


def _norm_ppf(q):
    from scipy.stats import norm
    return norm.ppf(q)


def test_oracle_sampled_mode_rates():
    text = open(__file__.replace('test_oracle.py', 'golden/circuits/surface_d3_r3.stim')).read()
    # Raise the noise so a small shot count has power.
    text = text.replace('0.001', '0.02')
    exact = fo.exact_detector_rates(text)
    shots = 40000
    dets, obs = fo.FrameOracle(text).run(shots, rng=np.random.default_rng(7))
    counts = np.concatenate([dets, obs], axis=1).sum(axis=0)
    z = _binomial_z(counts, shots, exact)
    assert np.max(np.abs(z)) < abs(_norm_ppf(1e-4 / (2 * len(z))))


def test_randomized_oracle_keeps_deterministic_detectors(golden):
    text = golden.circuit('surface_d5_r5')
    flat = st.parse_flat(text)
    # Strip noise: every detector is deterministic, so randomization must not show.
    keep = [i for i in flat if i.name not in ('X_ERROR', 'DEPOLARIZE1', 'DEPOLARIZE2')]
    dets, obs = fo.FrameOracle(keep, randomize=True).run(256, rng=np.random.default_rng(3))
    assert not dets.any() and not obs.any()


def test_randomized_oracle_exposes_random_detector():
    text = 'H 0\nM 0\nDETECTOR rec[-1]\nRX 1\nM 1\nDETECTOR rec[-1]\nR 2\nM 2\nDETECTOR rec[-1]\n'
    dets, _ = fo.FrameOracle(text, randomize=True).run(4096, rng=np.random.default_rng(1))
    rates = dets.mean(axis=0)
    assert abs(rates[0] - 0.5) < 0.05 and abs(rates[1] - 0.5) < 0.05 and rates[2] == 0


class _WrongCX(fo.FrameOracle):
    """Oracle with a planted mistake: CX forgets the control's Z update."""

    def _pair(self, n, ta, tb, x, z, rec):
        if n == 'CX' and ta.kind == st.QUBIT and tb.kind == st.QUBIT:
            x[tb.value] ^= x[ta.value]
            return
        return super()._pair(n, ta, tb, x, z, rec)


class _WrongDep2(fo.FrameOracle):
    """Oracle with a planted mistake: two-qubit fault codes applied to swapped qubits."""

    def _channel_codes_2q(self, k, n, a, npairs, S, fixed, fmap, rng):
        c = super()._channel_codes_2q(k, n, a, npairs, S, fixed, fmap, rng)
        return ((c & 3) << 2) | ((c >> 2) & 3)


@pytest.mark.parametrize('cls', [_WrongCX, _WrongDep2])
def test_fixtures_catch_planted_mistakes(golden, cls):
    fixtures = [f for f in golden.load('fault_vectors.json') if f['name'].startswith('surface')]
    bad, total = _mismatches(fixtures, cls)
    assert bad > 0


def test_fixture_faults_cover_every_channel_kind(golden):
    kinds = set()
    for f in golden.load('fault_vectors.json'):
        flat = st.parse_flat(f['circuit'])
        for s in f['sets']:
            for (k, j, c) in s['faults']:
                kinds.add(flat[k].name)
    for name in ['X_ERROR', 'Y_ERROR', 'Z_ERROR', 'DEPOLARIZE1', 'DEPOLARIZE2', 'PAULI_CHANNEL_1', 'PAULI_CHANNEL_2',
                 'E', 'ELSE_CORRELATED_ERROR', 'HERALDED_ERASE', 'HERALDED_PAULI_CHANNEL_1', 'MPP', 'MR', 'M']:
        assert name in kinds, name


# frame_simulator.test.cc:645-830 (classical_can_control_quantum, classical_controls): flips of the two
# results relative to the noiseless circuit, written as one detector per result.
CLASSICAL_CONTROL = [
    ('X_ERROR(1) 0\nM !0\nCX rec[-1] 1\nM 1\n', (1, 1)),
    ('X_ERROR(1) 0\nM !0\nCY rec[-1] 1\nM 1\n', (1, 1)),
    ('X_ERROR(1) 0\nM !0\nXCZ 1 rec[-1]\nM 1\n', (1, 1)),
    ('X_ERROR(1) 0\nM !0\nYCZ 1 rec[-1]\nM 1\n', (1, 1)),
    ('M 0\nCX rec[-1] 1\nM 1\n', (0, 0)),
    ('M !0\nCX rec[-1] 1\nM 1\n', (0, 0)),
    ('X_ERROR(1) 0\nM 0\nCX rec[-1] 1\nM 1\n', (1, 1)),
    ('X_ERROR(1) 0\nM 0\nH 1\nCZ rec[-1] 1\nH 1\nM 1\n', (1, 1)),
    ('X_ERROR(1) 0\nM 0\nCY rec[-1] 1\nM 1\n', (1, 1)),
]


@pytest.mark.parametrize('text,want', CLASSICAL_CONTROL)
def test_oracle_classical_control_known_answers(text, want):
    orc = fo.FrameOracle(text + 'DETECTOR rec[-2]\nDETECTOR rec[-1]\n')
    dets, _ = orc.run(16, rng=np.random.default_rng(0))
    assert (dets == np.array(want, dtype=bool)[None, :]).all()


def test_oracle_classical_control_copies_a_random_result():
    """frame_simulator.test.cc:726-741: X_ERROR(0.5) then CX rec[-1] 1: both results agree per shot."""
    orc = fo.FrameOracle('X_ERROR(0.5) 0\nM 0\nCX rec[-1] 1\nM 1\nDETECTOR rec[-2]\nDETECTOR rec[-1]\n')
    dets, _ = orc.run(1000, rng=np.random.default_rng(1))
    assert (dets[:, 0] == dets[:, 1]).all() and 400 < dets[:, 0].sum() < 600


@pytest.mark.parametrize('reset,zero,pair', [('RZ', 2, (0, 1)), ('RY', 1, (0, 2)), ('RX', 0, (1, 2))])
def test_randomized_oracle_observable_pauli_targets(reset, zero, pair):
    """frame_simulator.test.cc:1685-1743: after a reset, the observable on the reset basis's Pauli
    never flips and the other two are random and equal (shot by shot, which implies the reference's
    equal counts)."""
    text = f'{reset} 0\nOBSERVABLE_INCLUDE(0) X0\nOBSERVABLE_INCLUDE(1) Y0\nOBSERVABLE_INCLUDE(2) Z0\n'
    _, obs = fo.FrameOracle(text, randomize=True).run(1024, rng=np.random.default_rng(2))
    assert not obs[:, zero].any()
    a, b = pair
    assert (obs[:, a] == obs[:, b]).all() and 300 < obs[:, a].sum() < 700


def _zoo_counts_z(case, rates):
    return _binomial_z(case['counts'], case['shots'], rates)


def test_exact_rates_with_chains_heralds_and_pauli_channels_match_reference_counts(golden):
    """Gate-zoo circuits with PAULI_CHANNEL_1/2, heralded channels and correlated-error chains: the
    oracle's exact rates against the reference CLI's seeded counts (2^20 shots each)."""
    cases = golden.load('zoo_counts.json')
    assert len(cases) >= 6
    zmax = []
    for name, c in cases.items():
        exact = fo.exact_detector_rates(c['circuit'])
        z = _zoo_counts_z(c, exact)
        zcrit = abs(_norm_ppf(1e-4 / (2 * len(z))))
        assert np.max(np.abs(z)) < zcrit, (name, np.max(np.abs(z)))
        zmax.append(np.max(np.abs(z)))


def test_exact_rates_planted_chain_mistake_fails(golden):
    """Treating each ELSE_CORRELATED_ERROR as an independent channel (a dropped chain condition) must
    fail the same test on some circuit."""
    cases = golden.load('zoo_counts.json')
    orig = fo._channels

    def independent(flat):
        out = []
        for ch in orig(flat):
            if len(ch) > 1 and flat[ch[0][0]].name in ('E', 'CORRELATED_ERROR'):
                # every element at its own unconditional probability
                out.extend([[(k, j, c, flat[k].args[0])] for (k, j, c, _) in ch])
            else:
                out.append(ch)
        return out

    worst = 0.0
    try:
        fo._channels = independent
        for name, c in cases.items():
            z = _zoo_counts_z(c, fo.exact_detector_rates(c['circuit']))
            worst = max(worst, float(np.max(np.abs(z)) / abs(_norm_ppf(1e-4 / (2 * len(z))))))
    finally:
        fo._channels = orig
    assert worst > 1.0


@pytest.mark.parametrize('gate', ['PAULI_CHANNEL_2', 'HERALDED_PAULI_CHANNEL_1'])
def test_exact_rates_planted_outcome_order_mistake_fails(golden, gate):
    """Reversing the outcome order of one channel kind must fail the reference-count test."""
    case = golden.load('zoo_counts.json')['chain_stress']
    orig = fo.noise_sites

    def wrong(flat, allow_chains=False):
        out = []
        for (k, j, outs) in orig(flat, allow_chains):
            if flat[k].name == gate:
                outs = list(zip([c for c, _ in outs], [p for _, p in outs][::-1]))
            out.append((k, j, outs))
        return out

    try:
        fo.noise_sites = wrong
        z = _zoo_counts_z(case, fo.exact_detector_rates(case['circuit']))
    finally:
        fo.noise_sites = orig
    assert np.max(np.abs(z)) > abs(_norm_ppf(1e-4 / (2 * len(z))))
