"""Batched sampling (persistent kernel, HIP graph replay, autotuner) against separate plain calls.

A batch job must equal, bit for bit, the plain call `sample(shots, bit_packed=True,
separate_observables=True)` on a sampler with the same seed, at the same point of its stream."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _circuits():
    from stim_amd import gen
    out = []
    for (code, d, r, p) in [('surface_code:rotated_memory_z', 3, 3, 2e-3), ('surface_code:rotated_memory_z', 5, 5, 1e-3),
                            ('surface_code:rotated_memory_x', 7, 7, 4e-3), ('repetition_code:memory', 9, 9, 1e-2),
                            ('surface_code:unrotated_memory_z', 4, 2, 5e-3)]:
        out.append(gen.generate(code, distance=d, rounds=r, after_clifford_depolarization=p,
                                after_reset_flip_probability=p, before_measure_flip_probability=p))
    # A circuit with a random detector runs on the frame engine, as its own launch inside the batch.
    out.append('H 0\nM 0\nDETECTOR rec[-1]\nR 1\nX_ERROR(0.25) 1\nM 1\nDETECTOR rec[-1]\nOBSERVABLE_INCLUDE(0) rec[-1]\n')
    return out


SHOTS = [3000, 70001, 12345, 200000, 777, 4096]


def _samplers(seed0=100):
    import stim_amd
    out = [stim_amd.Circuit(t).compile_detector_sampler(seed=seed0 + k) for k, t in enumerate(_circuits())]
    for s in out:
        s.engine_info()   # builds the symptom tables, so small plain calls use them too
    return out


def _plain(s, shots):
    return s.sample(shots, bit_packed=True, separate_observables=True)


def _check(outputs, expected):
    for k, ((d, o), (ed, eo)) in enumerate(zip(outputs, expected)):
        np.testing.assert_array_equal(d.cpu().numpy(), ed, err_msg=f'job {k} detectors')
        np.testing.assert_array_equal(o.cpu().numpy(), eo, err_msg=f'job {k} observables')


def test_batch_equals_plain_calls():
    from stim_amd.batch import BatchSampler
    batch = BatchSampler(list(zip(_samplers(), SHOTS)))
    info = batch.info()
    assert info['jobs'] == 6 and info['kernel_jobs'] == 5 and info['separate_jobs'] == 1
    got = batch.run()
    _check(got, [_plain(s, n) for s, n in zip(_samplers(), SHOTS)])


def test_batch_graph_replays_continue_the_stream():
    """Graph replays with eager work in between; every replay equals the matching plain call."""
    from stim_amd.batch import BatchSampler
    batch = BatchSampler(list(zip(_samplers(), SHOTS)))
    ref = _samplers()
    for replay in range(3):
        got = batch.run(use_graph=True)
        _check(got, [_plain(s, n) for s, n in zip(ref, SHOTS)])


def test_one_workgroup_per_tile_schedule_equals_plain_calls():
    from stim_amd.batch import BatchSampler
    batch = BatchSampler(list(zip(_samplers(), SHOTS)))
    batch.set_schedule(False)
    ref = _samplers()
    for _ in range(2):
        _check(batch.run(use_graph=True), [_plain(s, n) for s, n in zip(ref, SHOTS)])
    assert batch.info()['persistent'] == 0


def test_same_sampler_twice_in_one_batch():
    import stim_amd
    from stim_amd.batch import BatchSampler
    text = _circuits()[1]
    s = stim_amd.Circuit(text).compile_detector_sampler(seed=9)
    batch = BatchSampler([(s, 5000), (s, 3000)])
    got = batch.run()
    ref = stim_amd.Circuit(text).compile_detector_sampler(seed=9)
    ref.engine_info()
    _check(got, [_plain(ref, 5000), _plain(ref, 3000)])
    # A plain call after the batch continues the stream.
    np.testing.assert_array_equal(_plain(s, 1000)[0], _plain(ref, 1000)[0])


def test_same_sampler_twice_graph_replays_continue_the_stream():
    """A sampler with two jobs in one batch advances by both jobs' tiles on every replay."""
    import stim_amd
    from stim_amd.batch import BatchSampler
    text = _circuits()[2]
    s = stim_amd.Circuit(text).compile_detector_sampler(seed=13)
    other = stim_amd.Circuit(_circuits()[0]).compile_detector_sampler(seed=14)
    batch = BatchSampler([(s, 6000), (other, 2000), (s, 3001)])
    ref = stim_amd.Circuit(text).compile_detector_sampler(seed=13)
    ref_o = stim_amd.Circuit(_circuits()[0]).compile_detector_sampler(seed=14)
    ref.engine_info()
    ref_o.engine_info()
    for _ in range(3):
        got = batch.run(use_graph=True)
        _check(got, [_plain(ref, 6000), _plain(ref_o, 2000), _plain(ref, 3001)])


def test_plain_call_between_runs_resyncs_the_batch():
    import stim_amd
    from stim_amd.batch import BatchSampler
    text = _circuits()[0]
    s = stim_amd.Circuit(text).compile_detector_sampler(seed=21)
    ref = stim_amd.Circuit(text).compile_detector_sampler(seed=21)
    ref.engine_info()
    batch = BatchSampler([(s, 4000)])
    _check(batch.run(use_graph=True), [_plain(ref, 4000)])
    np.testing.assert_array_equal(_plain(s, 999)[0], _plain(ref, 999)[0])
    _check(batch.run(use_graph=True), [_plain(ref, 4000)])


def test_autotuner_choice_keeps_batch_equal_to_plain_calls():
    from stim_amd.batch import BatchSampler
    batch = BatchSampler(list(zip(_samplers(), SHOTS)))
    kb, ms = batch.autotune(candidates=(8, 16, 32), iters=1)
    assert kb in (8, 16, 32) and ms > 0
    ref = _samplers()
    # Bring the reference samplers to the same stream position: the autotuner ran 2 runs per candidate
    # tile size and schedule (the schedule does not change the results).
    for s, n in zip(ref, SHOTS):
        for cand in (8, 16, 32):
            s.set_symptom_tile_kb(cand)
            for _ in range(4):
                _plain(s, n)
        s.set_symptom_tile_kb(kb)
    got = batch.run()
    _check(got, [_plain(s, n) for s, n in zip(ref, SHOTS)])
