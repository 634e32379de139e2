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
