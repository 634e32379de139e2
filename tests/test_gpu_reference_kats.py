"""Known-answer tests restated from the reference's own unit tests (expected outputs kept as data).

Each circuit is deterministic (its only noise has probability 1), so both engines must reproduce the
reference's expected bytes exactly.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ENGINES = ['frame', 'symptom']


def _sampler(text, engine, seed=3):
    import stim_amd
    s = stim_amd.Circuit(text).compile_detector_sampler(seed=seed)
    s.set_engine(engine)
    return s


@pytest.mark.parametrize('engine', ENGINES)
def test_certain_flip_fires_every_shot(engine):
    """frame_simulator_util.test.cc:31-44: X_ERROR(1) before M, one detector: '11111' over 5 shots."""
    s = _sampler('X_ERROR(1) 0\nM 0\nDETECTOR rec[-1]\n', engine)
    d = s.sample(5)
    assert d.shape == (5, 1) and d.dtype == np.bool_ and d.all()


@pytest.mark.parametrize('engine', ENGINES)
def test_dets_and_hits_text_with_observables(engine, tmp_path):
    """frame_simulator_util.test.cc:50-79: the dets format with observables prepended or appended,
    and hits with observables appended (observable 4 is column D + 4 = 6)."""
    text = 'X_ERROR(1) 0\nM 0 1\nDETECTOR rec[-1]\nDETECTOR rec[-2]\nOBSERVABLE_INCLUDE(4) rec[-2]\n'
    cases = [
        (dict(format='dets'), 'shot D1\nshot D1\n'),
        (dict(format='dets', prepend_observables=True), 'shot L4 D1\nshot L4 D1\n'),
        (dict(format='dets', append_observables=True), 'shot D1 L4\nshot D1 L4\n'),
        (dict(format='hits', append_observables=True), '1,6\n1,6\n'),
    ]
    for i, (kw, want) in enumerate(cases):
        s = _sampler(text, engine)
        path = tmp_path / f'out{i}.txt'
        s.sample_write(2, filepath=str(path), **kw)
        assert path.read_text() == want, kw


@pytest.mark.parametrize('engine', ENGINES)
def test_repeat_blocks_b8_detectors_and_observables(engine, tmp_path):
    """frame_simulator_util.test.cc:427-474 (obs_data): 1001 shots of b8 rows, 125 detector bytes per
    shot with exactly bits 0..398 set, and observable bytes 0x3."""
    text = '''REPEAT 399 {
    X_ERROR(1) 0
    MR 0
    DETECTOR rec[-1]
}
REPEAT 600 {
    MR 0
    DETECTOR rec[-1]
    OBSERVABLE_INCLUDE(0) rec[-1]
}
X_ERROR(1) 0
MR 0
OBSERVABLE_INCLUDE(0) rec[-1]
OBSERVABLE_INCLUDE(1) rec[-1]
MR 0
OBSERVABLE_INCLUDE(2) rec[-1]
'''
    s = _sampler(text, engine)
    dpath, opath = tmp_path / 'd.b8', tmp_path / 'o.b8'
    s.sample_write(1001, filepath=str(dpath), format='b8', obs_out_filepath=str(opath), obs_out_format='b8')
    dets = np.frombuffer(dpath.read_bytes(), dtype=np.uint8)
    obs = np.frombuffer(opath.read_bytes(), dtype=np.uint8)
    assert dets.size == 125 * 1001 and obs.size == 1001
    bits = np.unpackbits(dets.reshape(1001, 125), axis=1, bitorder='little')
    want = (np.arange(125 * 8) < 399).astype(np.uint8)
    assert (bits == want[None, :]).all()
    assert (obs == 0x3).all()


def _with_detectors(text, n):
    """The circuit plus one detector per measurement result, in result order."""
    return text + ''.join(f'DETECTOR rec[-{n - i}]\n' for i in range(n))


CHAINS = [
    ((0, 0, 0), '0000'), ((1, 0, 0), '1100'), ((0, 1, 0), '0110'), ((0, 0, 1), '0011'),
    ((1, 1, 0), '1100'), ((1, 1, 1), '1100'),
]


@pytest.mark.parametrize('probs,want', CHAINS)
def test_correlated_error_chains(probs, want):
    """frame_simulator.test.cc:425-545: an ELSE_CORRELATED_ERROR chain applies its first link that
    fires and skips the rest. Checked through the measurement sampler and, with one detector per
    result, through the detector sampler on both engines."""
    import stim_amd
    a, b, c = probs
    text = (f'CORRELATED_ERROR({a}) X0 X1\nELSE_CORRELATED_ERROR({b}) X1 X2\n'
            f'ELSE_CORRELATED_ERROR({c}) X2 X3\nM 0 1 2 3\n')
    expect = np.array([ch == '1' for ch in want])
    m = stim_amd.Circuit(text).compile_sampler(seed=1).sample(8)
    assert (m == expect[None, :]).all()
    for engine in ENGINES:
        d = _sampler(_with_detectors(text, 4), engine).sample(8)
        assert (d == expect[None, :]).all(), engine


def test_correlated_error_chain_then_new_chain():
    """frame_simulator.test.cc:535-545: a CORRELATED_ERROR after a chain starts a new chain."""
    import stim_amd
    text = ('CORRELATED_ERROR(1) X0 X1\nELSE_CORRELATED_ERROR(1) X1 X2\nELSE_CORRELATED_ERROR(1) X2 X3\n'
            'CORRELATED_ERROR(1) X3 X4\nM 0 1 2 3 4\n')
    expect = np.array([True, True, False, True, True])
    assert (stim_amd.Circuit(text).compile_sampler(seed=1).sample(8) == expect[None, :]).all()
    for engine in ENGINES:
        assert (_sampler(_with_detectors(text, 5), engine).sample(8) == expect[None, :]).all(), engine


def test_mpad_results():
    """frame_simulator.test.cc:1489-1500: MPAD 0 1 records 0 then 1 on every shot."""
    import stim_amd
    m = stim_amd.Circuit('MPAD 0 1\n').compile_sampler(seed=1).sample(100)
    assert m.shape == (100, 2) and not m[:, 0].any() and m[:, 1].all()


def test_pair_measurements_after_matching_resets():
    """frame_simulator.test.cc:1502-1518: MXX after RX, MYY after RY, MZZ after RZ all give 0."""
    import stim_amd
    text = 'RX 0 1\nMXX 0 1\nRY 0 1\nMYY 0 1\nRZ 0 1\nMZZ 0 1\n'
    m = stim_amd.Circuit(text).compile_sampler(seed=1).sample(100)
    assert m.shape == (100, 3) and not m.any()


def test_pair_measurement_inversions():
    """frame_simulator.test.cc:1520-1542: inverted targets of MXX/MYY/MZZ invert the result, and the
    three products of one pair multiply to -1 (result 8 = 1 ^ result 0 ^ result 4)."""
    import stim_amd
    text = 'MXX 0 1 0 !1 !0 1 !0 !1\nMYY 0 1 0 !1 !0 1 !0 !1\nMZZ 0 1 0 !1 !0 1 !0 !1\n'
    m = stim_amd.Circuit(text).compile_sampler(seed=1).sample(100)
    assert m.shape == (100, 12)
    for base in (0, 4, 8):
        assert (m[:, base + 1] == ~m[:, base]).all()
        assert (m[:, base + 2] == ~m[:, base]).all()
        assert (m[:, base + 3] == m[:, base]).all()
    assert (m[:, 8] == (True ^ m[:, 0] ^ m[:, 4])).all()
    # The X and Y products are random: both values occur.
    assert 0 < m[:, 0].sum() < 100 and 0 < m[:, 4].sum() < 100
