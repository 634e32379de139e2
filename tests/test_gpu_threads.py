"""Handles used from several threads at once.

Handles are not shared between threads (like the reference's sampler objects), but different handles
may run concurrently. The circuits below need different LDS per launch on both engines, so concurrent
launches also exercise the launchers' dynamic-LDS reservation. Seeded results depend only on the seed
and each handle's own sequence of calls, so every thread's rows must equal the rows of the same calls
made one after another.
"""

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

JOBS = [('surface_d3_r3', 'symptom'), ('surface_d5_r5', 'symptom'), ('color_d5_r3', 'frame'),
        ('surface_x_d5_r3', 'frame'), ('rep_d7_r5', 'symptom')]


def _run(text, engine, seed):
    import stim_amd
    s = stim_amd.Circuit(text).compile_detector_sampler(seed=seed)
    s.set_engine(engine)
    rows = [s.sample(3000 + 1000 * k, bit_packed=True, separate_observables=True) for k in range(3)]
    conv = stim_amd.Circuit(text).compile_m2d_converter()
    meas = stim_amd.Circuit(text).compile_sampler(seed=seed).sample(2000, bit_packed=True)
    rows.append(conv.convert(measurements=meas, bit_packed=True, append_observables=True))
    return rows


def test_handles_on_separate_threads(golden):
    texts = [golden.circuit(name) for name, _ in JOBS]
    expected = [_run(t, engine, 40 + i) for i, (t, (_, engine)) in enumerate(zip(texts, JOBS))]
    results = [None] * len(JOBS)
    errors = []

    def worker(i):
        try:
            results[i] = _run(texts[i], JOBS[i][1], 40 + i)
        except Exception as e:  # reported below with the job's name
            errors.append((JOBS[i][0], repr(e)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(JOBS))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for (name, _), want, got in zip(JOBS, expected, results):
        assert got is not None, name
        for w, g in zip(want, got):
            if isinstance(w, tuple):
                for a, b in zip(w, g):
                    assert np.array_equal(a, b), name
            else:
                assert np.array_equal(w, g), name
