"""GPU parity: the HIP frame kernel (through the C ABI) against the reference's own outputs and the oracle.

Fixed-fault mode must be bit-exact; seeded mode must pass binomial tests against exact rates.
"""

import os
import random
import tempfile

import numpy as np
import pytest

from oracle import frame_oracle as fo
from oracle import stim_text as st
import zoo

pytestmark = pytest.mark.gpu


ENGINES = ['frame', 'symptom']


def _sampler(text, seed=1, engine='auto'):
    import stim_amd
    s = stim_amd.Circuit(text).compile_detector_sampler(seed=seed)
    s.set_engine(engine)
    return s


def _rows(dets, obs):
    return [''.join('1' if b else '0' for b in np.concatenate([dets[i], obs[i]])) for i in range(len(dets))]


# ---------------------------------------------------------------- fixed-fault parity (bit-exact)

@pytest.mark.parametrize('engine', ENGINES)
def test_fixed_faults_match_reference_vectors(golden, engine):
    fixtures = golden.load('fault_vectors.json')
    for f in fixtures:
        print('circuit', f['name'], engine, flush=True)
        s = _sampler(f['circuit'], engine=engine)
        faults = [[tuple(x) for x in st_['faults']] for st_ in f['sets']]
        dets, obs = s.sample_with_faults(faults, separate_observables=True)
        got = _rows(dets, obs)
        exp = [st_['expected'] for st_ in f['sets']]
        bad = [i for i in range(len(exp)) if got[i] != exp[i]]
        assert not bad, (f['name'], bad[:5], [f['sets'][i]['faults'] for i in bad[:3]])


def _random_fault_sets(flat, n_shots, rng, max_faults=6):
    chans = fo.noise_sites(flat, allow_chains=True)
    out = []
    for _ in range(n_shots):
        k = rng.randint(0, max_faults)
        picks = rng.sample(chans, k=min(k, len(chans)))
        out.append([(a, b, rng.choice(outs)[0]) for (a, b, outs) in picks])
    return out


@pytest.mark.parametrize('engine', ENGINES)
@pytest.mark.parametrize('seed', range(6))
def test_fixed_faults_match_oracle_on_gate_zoo(seed, engine):
    text = zoo.gate_zoo(100 + seed, n_qubits=12 + 4 * seed, layers=14)
    flat = st.parse_flat(text)
    rng = random.Random(seed)
    faults = _random_fault_sets(flat, 300, rng)
    print('zoo seed', seed, engine, flush=True)
    s = _sampler(text, engine=engine)
    dets, obs = s.sample_with_faults(faults, separate_observables=True)
    edets, eobs = fo.FrameOracle(flat).run(len(faults), faults=faults)
    np.testing.assert_array_equal(dets, edets)
    np.testing.assert_array_equal(obs, eobs)


@pytest.mark.parametrize('engine', ENGINES)
@pytest.mark.parametrize('name', ['surface_d5_r5', 'color_d5_r3', 'surface_x_d5_r3'])
def test_fixed_faults_bit_packed_match_oracle(golden, name, engine):
    text = golden.circuit(name)
    flat = st.parse_flat(text)
    faults = _random_fault_sets(flat, 700, random.Random(11), max_faults=12)
    s = _sampler(text, engine=engine)
    dets, obs = s.sample_with_faults(faults, separate_observables=True, bit_packed=True)
    edets, eobs = fo.FrameOracle(flat).run(len(faults), faults=faults)
    np.testing.assert_array_equal(dets, fo.pack_bits(edets))
    np.testing.assert_array_equal(obs, fo.pack_bits(eobs))


@pytest.mark.parametrize('engine', ENGINES)
def test_fixed_faults_d25_full_size_match_oracle(golden, engine):
    text = golden.circuit('surface_d25_r25')
    flat = st.parse_flat(text)
    faults = _random_fault_sets(flat, 64, random.Random(5), max_faults=40)
    s = _sampler(text, engine=engine)
    dets, obs = s.sample_with_faults(faults, separate_observables=True, bit_packed=True)
    edets, eobs = fo.FrameOracle(flat).run(len(faults), faults=faults)
    np.testing.assert_array_equal(dets, fo.pack_bits(edets))
    np.testing.assert_array_equal(obs, fo.pack_bits(eobs))


def test_xor_linearity_of_fault_injection(golden):
    """Detection events of a fault set equal the XOR of its single faults' events."""
    text = golden.circuit('surface_d5_r5')
    flat = st.parse_flat(text)
    rng = random.Random(3)
    sets = _random_fault_sets(flat, 50, rng, max_faults=5)
    singles = [[f] for fs in sets for f in fs]
    s = _sampler(text)
    d_sets = s.sample_with_faults(sets, bit_packed=True, append_observables=True)
    d_single = s.sample_with_faults(singles, bit_packed=True, append_observables=True)
    pos = 0
    for i, fs in enumerate(sets):
        acc = np.zeros(d_sets.shape[1], dtype=np.uint8)
        for _ in fs:
            acc ^= d_single[pos]
            pos += 1
        np.testing.assert_array_equal(d_sets[i], acc)


@pytest.mark.parametrize('engine', ENGINES)
def test_noiseless_circuits_give_zero_detectors(golden, engine):
    for name in ['surface_d5_r5', 'surface_d25_r25', 'color_d5_r3', 'rep_d7_r5']:
        text = golden.circuit(name)
        for p in ['0.001', '0.002']:
            text = text.replace(f'({p})', '(0)')
        s = _sampler(text, engine=engine)
        d, o = s.sample(5000, bit_packed=True, separate_observables=True)
        assert not d.any() and not o.any(), name


# ---------------------------------------------------------------- reference API vectors

KAT = """
X_ERROR(1) 0
M 0 1 2
DETECTOR rec[-3] rec[-2]
DETECTOR rec[-3] rec[-1]
DETECTOR rec[-1] rec[-2]
OBSERVABLE_INCLUDE(0) rec[-3]
OBSERVABLE_INCLUDE(3) rec[-2] rec[-1]
OBSERVABLE_INCLUDE(0) rec[-2]
"""


@pytest.mark.parametrize('engine', ENGINES)
def test_reference_sample_vectors(engine):
    """Vectors from the reference's compiled_detector_sampler_pybind_test.py:37-160."""
    s = _sampler(KAT, engine=engine)
    np.testing.assert_array_equal(s.sample(5), np.array([[1, 1, 0]] * 5, dtype=np.bool_))
    np.testing.assert_array_equal(s.sample(5, prepend_observables=True),
                                  np.array([[1, 0, 0, 0, 1, 1, 0]] * 5, dtype=np.bool_))
    np.testing.assert_array_equal(s.sample(5, append_observables=True),
                                  np.array([[1, 1, 0, 1, 0, 0, 0]] * 5, dtype=np.bool_))
    dets, obs = s.sample(5, separate_observables=True)
    np.testing.assert_array_equal(dets, np.array([[1, 1, 0]] * 5, dtype=np.bool_))
    np.testing.assert_array_equal(obs, np.array([[1, 0, 0, 0]] * 5, dtype=np.bool_))
    np.testing.assert_array_equal(s.sample(5, bit_packed=True), np.array([[0b011]] * 5, dtype=np.uint8))
    np.testing.assert_array_equal(s.sample(5, prepend_observables=True, bit_packed=True),
                                  np.array([[0b0110001]] * 5, dtype=np.uint8))
    np.testing.assert_array_equal(s.sample(5, append_observables=True, bit_packed=True),
                                  np.array([[0b0001011]] * 5, dtype=np.uint8))
    dets, obs = s.sample(5, separate_observables=True, bit_packed=True)
    np.testing.assert_array_equal(dets, np.array([[0b011]] * 5, dtype=np.uint8))
    np.testing.assert_array_equal(obs, np.array([[0b0001]] * 5, dtype=np.uint8))
    t = _sampler('X_ERROR(1) 0\nM 0\nDETECTOR rec[-1]\n', engine=engine)
    np.testing.assert_array_equal(t.sample(shots=2), np.array([[1], [1]], dtype=np.bool_))


def test_sample_argument_errors():
    s = _sampler(KAT)
    with pytest.raises(ValueError):
        s.sample(3, separate_observables=True, append_observables=True)
    with pytest.raises(ValueError):
        s.sample(3, dets_out=np.zeros((3, 3), dtype=np.uint8))
    with pytest.raises(ValueError):
        s.sample(3, bit_packed=True, dets_out=np.zeros((3, 2), dtype=np.uint8))
    with pytest.raises(ValueError):
        s.sample(3, bit_packed=True, dets_out=np.zeros((3,), dtype=np.uint8))


@pytest.mark.parametrize('engine', ENGINES)
def test_manual_output_buffers(engine):
    s = _sampler(KAT, engine=engine)
    dets = np.zeros((11, 1), dtype=np.uint8)
    obs = np.zeros((11, 1), dtype=np.uint8)
    d2, o2 = s.sample(11, separate_observables=True, bit_packed=True, dets_out=dets, obs_out=obs)
    assert d2 is dets and o2 is obs
    assert (dets == 0b011).all() and (obs == 0b0001).all()
    dets = np.zeros((11, 3), dtype=np.bool_)
    s.sample(11, dets_out=dets)
    assert dets[:, :2].all() and not dets[:, 2].any()


@pytest.mark.parametrize('engine', ENGINES)
def test_sample_write_formats_match_reference_bytes(golden, engine):
    fx = golden.load('formats.json')
    s = _sampler(fx['circuit'], engine=engine)
    with tempfile.TemporaryDirectory() as d:
        for key, hexdata in sorted(fx['outputs'].items()):
            fmt, mode, shots = key.split('/')
            shots = int(shots)
            path = os.path.join(d, 'out')
            obs_path = os.path.join(d, 'obs')
            if mode in ('plain', 'append', 'prepend'):
                # The CLI turns on prepend_observables for the dets format unless observables are
                # appended (src/stim/cmd/command_detect.cc:45-47); sample_write itself does not.
                prepend = mode == 'prepend' or (fmt == 'dets' and mode == 'plain')
                s.sample_write(shots, filepath=path, format=fmt, prepend_observables=prepend,
                               append_observables=mode == 'append')
                got = open(path, 'rb').read()
            else:
                s.sample_write(shots, filepath=path, format=fmt, obs_out_filepath=obs_path, obs_out_format=fmt)
                got = open(path if mode == 'obs_out_dets' else obs_path, 'rb').read()
            assert got.hex() == hexdata, key


def test_sample_write_conflicting_observable_flags():
    s = _sampler(KAT)
    with tempfile.TemporaryDirectory() as d:
        with pytest.raises(IndexError):
            s.sample_write(3, filepath=os.path.join(d, 'a'), format='b8', prepend_observables=True,
                           append_observables=True)
        with pytest.raises(ValueError):
            s.sample_write(3, filepath=os.path.join(d, 'a'), format='nope')


# ---------------------------------------------------------------- seeded statistics

def _z(counts, shots, p):
    p = np.clip(np.asarray(p, dtype=float), 1e-300, 1)
    return (np.asarray(counts, dtype=float) - shots * p) / np.sqrt(shots * p * (1 - p))


def _zcrit(n, alpha=1e-4):
    from scipy.stats import norm
    return abs(norm.ppf(alpha / (2 * n)))


@pytest.mark.parametrize('engine', ENGINES)
@pytest.mark.parametrize('name', ['surface_d5_r5', 'surface_d3_r3', 'color_d5_r3'])
def test_seeded_rates_match_exact_rates(golden, name, engine):
    """Per-detector firing rates pass a Bonferroni binomial test (alpha=1e-4) against exact rates."""
    rates = np.array(golden.load('dem_rates.json')[name])
    s = _sampler(golden.circuit(name), seed=12345, engine=engine)
    shots = 1 << 21
    d = s.sample(shots, bit_packed=True, append_observables=True)
    n = len(rates)
    bits = np.unpackbits(d, axis=1, bitorder='little')[:, :n]
    counts = bits.sum(axis=0)
    z = _z(counts, shots, rates)
    assert np.max(np.abs(z)) < _zcrit(n), np.argsort(-np.abs(z))[:5]
    # Global chi-square over all detectors.
    from scipy.stats import chi2
    stat = float((z ** 2).sum())
    assert stat < chi2.ppf(1 - 1e-4, df=n)


@pytest.mark.parametrize('engine', ENGINES)
def test_seeded_pair_correlations_match_exact(golden, engine):
    """Per-mechanism check: co-firing rate of detector pairs joined by single faults."""
    name = 'surface_d3_r3'
    text = golden.circuit(name).replace('0.001', '0.01')
    flat = st.parse_flat(text)
    chans = fo.noise_sites(flat)
    orc = fo.FrameOracle(flat)
    items = [(k, j, c, p) for (k, j, outs) in chans for (c, p) in outs]
    dets, obs = orc.run(len(items), faults=[[(k, j, c)] for (k, j, c, _) in items])
    sym = np.concatenate([dets, obs], axis=1)
    # Exact P(d1 and d2 both fire) by enumerating channels is expensive; instead compare the XOR
    # rate of pairs: P(d1 ^ d2) is an exact single-bit parity rate computable like a detector rate.
    n = sym.shape[1]
    pairs = [(a, b) for a in range(n) for b in range(a + 1, n) if (sym[:, a] & sym[:, b]).any()]
    assert pairs
    logs = np.zeros(len(pairs))
    pos = 0
    for (k, j, outs) in chans:
        q = np.zeros(len(pairs))
        for (c, p) in outs:
            row = sym[pos]
            q += p * np.array([row[a] ^ row[b] for a, b in pairs])
            pos += 1
        logs += np.log(np.abs(1 - 2 * q))
    exact = (1 - np.exp(logs)) / 2
    s = _sampler(text, seed=99, engine=engine)
    shots = 1 << 20
    d = s.sample(shots, append_observables=True)
    counts = np.array([(d[:, a] ^ d[:, b]).sum() for a, b in pairs])
    z = _z(counts, shots, exact)
    assert np.max(np.abs(z)) < _zcrit(len(pairs))


def _column_counts(packed, ncols, chunk=1 << 17):
    """Set-bit count of every column of bit-packed little-endian rows, a chunk of rows and a bit plane
    at a time."""
    counts = np.zeros(packed.shape[1] * 8, dtype=np.int64)
    for r0 in range(0, len(packed), chunk):
        part = packed[r0:r0 + chunk]
        for b in range(8):
            counts[b::8] += ((part >> b) & 1).sum(axis=0, dtype=np.int64)
    return counts[:ncols]


def _d25_exact_rates(golden):
    import gzip
    from oracle import dem_oracle as do
    with gzip.open(os.path.join(golden.dir, 'dems', 'surface_d25_r25_folded.dem.gz'), 'rt') as f:
        model = do.parse(f.read())
    assert (model[0], model[1]) == (15600, 1)
    return do.exact_rates(model)


@pytest.mark.parametrize('engine', ENGINES)
def test_d25_full_size_rates_and_shape(golden, engine):
    """The headline circuit (d=25, 25 rounds, p=1e-3) at its full 2^20 shots on each engine: every
    detector's and the observable's rate against the exact rate of the reference's own error model of
    this circuit (stim analyze_errors, committed), by per-column Bonferroni z-tests and a global
    chi-square; padding bits stay zero; the same seed reproduces."""
    text = golden.circuit('surface_d25_r25')
    s = _sampler(text, seed=2, engine=engine)
    shots = 1 << 20
    dets, obs = s.sample(shots, bit_packed=True, separate_observables=True)
    assert dets.shape == (shots, 1950) and obs.shape == (shots, 1)
    assert not (obs >> 1).any()
    exact = _d25_exact_rates(golden)
    counts = np.concatenate([_column_counts(dets, 15600), _column_counts(obs, 1)])
    z = _z(counts, shots, exact)
    assert np.max(np.abs(z)) < _zcrit(len(z)), np.argsort(-np.abs(z))[:5]
    # Chi-square over all columns: the mean of z^2 is 1 for independent columns; neighbouring
    # detectors are correlated, which widens its spread, so the bound is loose (a 5% rate bias on every
    # detector gives a mean of ~45).
    assert abs(float(np.mean(z * z)) - 1.0) < 0.1, float(np.mean(z * z))
    # Different calls draw different randomness; same seed reproduces.
    again = _sampler(text, seed=2, engine=engine).sample(1024, bit_packed=True, separate_observables=True)[0]
    np.testing.assert_array_equal(again, dets[:1024])


@pytest.mark.parametrize('engine', ENGINES)
def test_consecutive_calls_draw_fresh_randomness(engine):
    """Each call continues the sampler's random stream: no row block of one call reappears in the
    next (a reused stream would repeat whole tiles of rows), and a call split in two has the rate of
    one call."""
    from stim_amd import gen
    text = gen.generate('surface_code:rotated_memory_z', distance=5, rounds=5, after_clifford_depolarization=0.02,
                        after_reset_flip_probability=0.02, before_measure_flip_probability=0.02)
    s = _sampler(text, seed=11, engine=engine)
    calls = [s.sample(8192, bit_packed=True) for _ in range(3)]
    weight = [np.unpackbits(c, axis=1).sum(axis=1) for c in calls]
    seen = set()
    for c, w in zip(calls[:2], weight[:2]):
        seen.update(bytes(r) for r, k in zip(c, w) if k >= 12)
    heavy = [bytes(r) for r, k in zip(calls[2], weight[2]) if k >= 12]
    assert len(heavy) > 1000
    repeats = sum(r in seen for r in heavy)
    assert repeats <= 2, repeats


def test_wide_symptom_records_agree_with_frame_engine():
    """More than 2^15 detectors: the symptom engine uses its 8-byte records; rates must agree with the
    frame engine's."""
    from stim_amd import gen
    text = gen.generate('repetition_code:memory', distance=185, rounds=185, after_clifford_depolarization=0.01,
                        after_reset_flip_probability=0.01, before_measure_flip_probability=0.01)
    shots = 1 << 15
    a = _sampler(text, seed=31, engine='frame').sample(shots, bit_packed=True, append_observables=True)
    b = _sampler(text, seed=32, engine='symptom').sample(shots, bit_packed=True, append_observables=True)
    n = 184 * 186 + 1
    assert n > 1 << 15
    ra = np.unpackbits(a, axis=1, bitorder='little')[:, :n].mean(axis=0)
    rb = np.unpackbits(b, axis=1, bitorder='little')[:, :n].mean(axis=0)
    p = (ra + rb) / 2
    sd = np.sqrt(np.maximum(p * (1 - p), 1e-12) * 2 / shots)
    assert np.max(np.abs((ra - rb) / sd)) < _zcrit(n)


def test_engines_agree_statistically_on_zoo():
    """Frame and symptom engines sample the same distribution on an all-gate circuit."""
    text = zoo.gate_zoo(7, n_qubits=10, layers=12)
    shots = 1 << 18
    a = _sampler(text, seed=3, engine='frame').sample(shots, append_observables=True)
    b = _sampler(text, seed=4, engine='symptom').sample(shots, append_observables=True)
    ra, rb = a.mean(axis=0), b.mean(axis=0)
    p = (ra + rb) / 2
    sd = np.sqrt(np.maximum(p * (1 - p), 1e-12) * 2 / shots)
    z = (ra - rb) / sd
    assert np.max(np.abs(z)) < _zcrit(len(z))


def test_random_clifford_engines_agree():
    """The random-Clifford stress family (noisy MPP products, feedback, PAULI_CHANNEL_1, DEPOLARIZE2):
    both engines sample the same detector rates."""
    from stim_amd import gen
    text = gen.random_clifford(n_qubits=400, layers=30, seed=3, px=2e-3, py=2e-3, pz=2e-3, p2=5e-3, n_products=60,
                               mpp_flip=0.01)
    shots = 1 << 16
    a = _sampler(text, seed=41, engine='frame').sample(shots, append_observables=True)
    b = _sampler(text, seed=42, engine='symptom').sample(shots, append_observables=True)
    ra, rb = a.mean(axis=0), b.mean(axis=0)
    assert ra.mean() > 0.005
    p = (ra + rb) / 2
    sd = np.sqrt(np.maximum(p * (1 - p), 1e-12) * 2 / shots)
    assert np.max(np.abs((ra - rb) / sd)) < _zcrit(len(ra))


def test_auto_engine_is_symptom_for_deterministic_circuits(golden):
    s = _sampler(golden.circuit('surface_d5_r5'))
    info = s.engine_info()
    assert info['engine'] == 'symptom'
    assert info['columns'] > 0 and info['entries'] > 0


def test_random_detectors_use_frame_engine_and_are_fair():
    """A detector on a random measurement is uniform, as with the reference's frame randomization."""
    text = 'H 0\nM 0\nDETECTOR rec[-1]\nR 1\nX_ERROR(0.25) 1\nM 1\nDETECTOR rec[-1]\n'
    s = _sampler(text, seed=8)
    assert s.engine_info()['engine'] == 'frame'
    d = s.sample(1 << 16)
    z = _z(d.sum(axis=0), 1 << 16, np.array([0.5, 0.25]))
    assert np.max(np.abs(z)) < _zcrit(2)
    with pytest.raises(ValueError):
        s.set_engine('symptom')


@pytest.mark.parametrize('engine', ENGINES)
def test_rates_match_exact_rates_on_channel_zoo(golden, engine):
    """Circuits with PAULI_CHANNEL_1/2, heralded channels and correlated-error chains (including a
    high-probability stress circuit): each engine's per-detector rates against the oracle's exact
    rates, which the reference CLI's seeded counts pin (zoo_counts.json)."""
    cases = golden.load('zoo_counts.json')
    for name, c in cases.items():
        exact = fo.exact_detector_rates(c['circuit'])
        s = _sampler(c['circuit'], seed=77, engine=engine)
        if engine == 'symptom':
            assert s.engine_info()['engine'] == 'symptom'
        shots = 1 << 21
        d = s.sample(shots, bit_packed=True, append_observables=True)
        counts = _column_counts(d, len(exact))
        z = _z(counts, shots, exact)
        assert np.max(np.abs(z)) < _zcrit(len(z)), (name, np.argsort(-np.abs(z))[:5])


OVERFLOW_CIRCUIT = '''R 0 1 2
DEPOLARIZE2(0.2) 0 1
DEPOLARIZE1(0.2) 2
X_ERROR(0.2) 0
M 0 1 2
DETECTOR rec[-3]
DETECTOR rec[-3]
DETECTOR rec[-3]
DETECTOR rec[-2]
DETECTOR rec[-2]
DETECTOR rec[-2]
DETECTOR rec[-1] rec[-3]
OBSERVABLE_INCLUDE(0) rec[-2]
OBSERVABLE_INCLUDE(1) rec[-3]
'''


@pytest.mark.parametrize('name', ['surface_d5_r5', 'color_d5_r3', 'surface_x_d5_r3', 'rep_d7_r5', 'surface_d25_r25',
                                  'overflow'])
def test_specialised_scan_kernel_equals_generic(golden, name):
    """The specialised scan kernel (one outcome record per site and outcome) draws the generic kernel's
    hits, so rows are bit-identical in every output layout, including outcomes of more than four
    detectors (overflow lists) and both observables."""
    text = OVERFLOW_CIRCUIT if name == 'overflow' else golden.circuit(name)
    shots = (1 << 14) if name == 'surface_d25_r25' else (1 << 16)
    for kw in [dict(separate_observables=True), dict(append_observables=True), dict(prepend_observables=True),
               dict()]:
        a = _sampler(text, seed=31, engine='symptom')
        b = _sampler(text, seed=31, engine='symptom')
        b.set_scan_variant('generic')
        for packed in (True, False):
            ra = a.sample(shots, bit_packed=packed, **kw)
            rb = b.sample(shots, bit_packed=packed, **kw)
            if isinstance(ra, tuple):
                for x, y in zip(ra, rb):
                    np.testing.assert_array_equal(x, y, err_msg=f'{name} {kw} packed={packed}')
            else:
                np.testing.assert_array_equal(ra, rb, err_msg=f'{name} {kw} packed={packed}')
            assert (ra[0] if isinstance(ra, tuple) else ra).any()


def test_outcome_blocks_repeat_across_rounds(golden):
    """The specialised kernel stores each distinct block of a site's outcome records (relative to the
    site's lowest detector) once: the d=25, 25-round memory circuit's 108,674 sites and 979,874
    outcomes reduce to the blocks of a few rounds, small enough to stay in L2."""
    s = _sampler(golden.circuit('surface_d25_r25'), engine='symptom')
    info = s.engine_info()
    assert info['sites'] == 108674
    assert 0 < info['outcome_records'] < 979874 // 5, info
    assert info['outcome_blocks'] < info['sites'] // 5, info
