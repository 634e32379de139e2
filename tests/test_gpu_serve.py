"""Serving: admission, continuous batching, and results that do not depend on the mix.

Pinned by: the exact per-detector rates from the reference's detector error model (dem_rates.json)
for a served request, and bit-equality of one request run alone and inside a mix of others."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gen(code, d, r, p):
    from stim_amd import gen
    return gen.generate(code, distance=d, rounds=r, after_clifford_depolarization=p, after_reset_flip_probability=p,
                        before_measure_flip_probability=p)


def _sampler(text, seed):
    import stim_amd
    return stim_amd.Circuit(text).compile_detector_sampler(seed=seed)


def _run_until(srv, rids):
    for _ in range(100000):
        if all(srv.done(r) for r in rids):
            return
        srv.step()
    raise AssertionError('requests did not finish')


def test_request_result_is_the_same_alone_and_in_a_mix():
    from stim_amd.serve import Server
    x = _sampler(_gen('surface_code:rotated_memory_z', 5, 5, 2e-3), seed=3)
    alone = Server()
    r = alone.submit(x, 50000, seed=7)
    _run_until(alone, [r])
    a = alone.result(r)
    # In a mix, with small steps so the request is split over many steps and shares each with others.
    others = [_sampler(_gen('surface_code:rotated_memory_z', 7, 7, 1e-3), seed=4),
              _sampler(_gen('repetition_code:memory', 15, 15, 5e-3), seed=5),
              _sampler(_gen('surface_code:rotated_memory_x', 3, 3, 8e-3), seed=6)]
    mix = Server(step_cost_bytes=3e5)
    rids = [mix.submit(others[0], 120000, seed=1), mix.submit(others[1], 3000, seed=2)]
    mix.step()
    target = mix.submit(x, 50000, seed=7)
    rids += [target, mix.submit(others[2], 90000, seed=3), mix.submit(x, 50000, seed=8)]
    _run_until(mix, rids)
    b = mix.result(target)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert mix.stats()['steps'] > 5
    # A different request seed draws a different stream.
    c = mix.result(rids[-1])
    assert not np.array_equal(c[0], b[0])


def test_served_rates_match_exact_rates(golden):
    from scipy.stats import norm
    from stim_amd.serve import Server
    rates = np.array(golden.load('dem_rates.json')['surface_d3_r3'])
    s = _sampler(golden.circuit('surface_d3_r3'), seed=9)
    srv = Server(step_cost_bytes=1e6)
    shots = 1 << 20
    r = srv.submit(s, shots, seed=11)
    _run_until(srv, [r])
    dets, obs = srv.result(r)
    D = s.num_detectors
    bits = np.unpackbits(dets, axis=1, bitorder='little')[:, :D]
    counts = np.concatenate([bits.sum(axis=0), np.unpackbits(obs, axis=1, bitorder='little')[:, :1].sum(axis=0)])
    p = np.clip(rates[:len(counts)], 1e-300, 1)
    z = (counts - shots * p) / np.sqrt(shots * p * (1 - p))
    assert np.max(np.abs(z)) < abs(norm.ppf(1e-4 / (2 * len(z))))


def test_admission_waits_for_the_memory_budget():
    from stim_amd.serve import Server
    s = _sampler(_gen('surface_code:rotated_memory_z', 5, 5, 1e-3), seed=12)
    # 16 B/shot: 40000 shots need ~640 KB each; a 1 MB budget holds one at a time.
    srv = Server(budget_bytes=1 << 20)
    r1 = srv.submit(s, 40000, seed=1)
    r2 = srv.submit(s, 40000, seed=2)
    srv.step()
    st = srv.stats()
    assert st['active'] + (1 if srv.done(r1) else 0) == 1 and st['queued'] == 1
    _run_until(srv, [r1])
    srv.step()
    assert srv.stats()['queued'] == 1          # r1's rows are still held
    srv.release(r1)
    _run_until(srv, [r2])
    assert srv.stats()['peak_bytes_in_use'] <= 1 << 20
    with pytest.raises(ValueError):
        srv.submit(s, 10_000_000, seed=3)      # larger than the whole budget


def test_replay_trace_completes_every_request():
    from stim_amd.serve import Server
    samplers = [_sampler(_gen('surface_code:rotated_memory_z', d, d, 1e-3), seed=d) for d in (3, 5, 7)]
    rng = np.random.default_rng(1)
    n = 40
    arrivals = np.sort(rng.uniform(0, 0.2, n))
    chosen = [samplers[k] for k in rng.integers(0, 3, n)]
    shots = [int(x) for x in np.exp(rng.uniform(np.log(1e3), np.log(2e5), n))]
    srv = Server()
    lat, wall = srv.replay(chosen, shots, list(range(n)), arrivals)
    assert np.all(lat > 0) and wall >= arrivals[-1]
    st = srv.stats()
    assert st['queued'] == 0 and st['active'] == 0 and st['bytes_in_use'] == 0
