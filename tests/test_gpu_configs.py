"""The benchmark configurations at their real shapes on the GPU, beyond the headline circuit.

- 10,000-qubit, 1,000-layer random-Clifford circuit, 2^16 shots (frame engine): fixed faults bit-exact
  against the oracle's measurement flips (clifford_faults.npz, written by make_clifford_faults.py),
  and detection events (back-to-back product comparisons) all zero, noiseless or sampled.
- 64 heterogeneous surface codes in one persistent kernel: every job equal bit for bit to its plain
  call, and the small-distance jobs' per-detector rates against the oracle's exact rates.
- The 200-circuit serving pool: a mixed request stream through the server; requests re-served alone
  are bit-identical to their results in the mix, and served rates match exact rates.
"""

import os
import sys

import numpy as np
import pytest

from oracle import frame_oracle as fo

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _z(counts, shots, p):
    p = np.clip(np.asarray(p, dtype=float), 1e-300, 1)
    return (np.asarray(counts, dtype=float) - shots * p) / np.sqrt(shots * p * (1 - p))


def _zcrit(n, alpha=1e-4):
    from scipy.stats import norm
    return abs(norm.ppf(alpha / (2 * n)))


def _device_column_counts(t, ncols):
    """Per-column set-bit counts of bit-packed device rows (torch uint8, shape (shots, bytes))."""
    import torch
    counts = torch.zeros(t.shape[1] * 8, dtype=torch.int64, device=t.device)
    for b in range(8):
        counts[b::8] = ((t >> b) & 1).sum(dim=0, dtype=torch.int64)
    return counts[:ncols].cpu().numpy()


# ---------------------------------------------------------------- random Clifford, 10,000 x 1,000

@pytest.fixture(scope='module')
def clifford_text():
    from stim_amd import gen
    return gen.random_clifford(n_qubits=10000, layers=1000, seed=2026)


def test_clifford_full_size_fixed_faults_match_oracle(clifford_text):
    import stim_amd
    fx = np.load(os.path.join(ROOT, 'tests', 'golden', 'clifford_faults.npz'))
    faults = [[tuple(int(v) for v in f) for f in shot] for shot in fx['faults']]
    s = stim_amd.Circuit(clifford_text).compile_sampler(skip_reference_sample=True, seed=3)
    assert s.num_measurements == int(fx['num_results'])
    got = s.sample_with_faults(faults, bit_packed=True)
    np.testing.assert_array_equal(got, fx['rows'])
    assert fx['rows'].any()


def test_clifford_full_size_detectors(clifford_text):
    """Every detector compares two back-to-back measurements of one product, so the rows are zero
    whatever the noise: sampled at the full 2^16 shots on the frame engine, and with faults."""
    import stim_amd
    s = stim_amd.Circuit(clifford_text).compile_detector_sampler(seed=5)
    assert s.num_detectors == 50000 and s.num_observables == 1
    # 2^16 shots against ~4*10^7 symptom columns: the automatic choice keeps the frame engine (the
    # symptom table would cost more to build than the call).
    d, o = s.sample(1 << 16, bit_packed=True, separate_observables=True)
    assert d.shape == (1 << 16, 6250) and not d.any() and not o.any()
    fx = np.load(os.path.join(ROOT, 'tests', 'golden', 'clifford_faults.npz'))
    faults = [[tuple(int(v) for v in f) for f in shot] for shot in fx['faults']]
    d, o = s.sample_with_faults(faults, bit_packed=True, separate_observables=True)
    assert not d.any() and not o.any()


# ---------------------------------------------------------------- batch of 64 surface codes

def test_batch64_configuration_equals_plain_calls_and_exact_rates():
    import torch
    import stim_amd
    from stim_amd import _native as N
    from stim_amd import gen
    from stim_amd.batch import BatchSampler
    import bench
    cfgs = bench.batch64_configs()
    texts = [gen.generate('surface_code:rotated_memory_z', distance=d, rounds=d, after_clifford_depolarization=p,
                          after_reset_flip_probability=p, before_measure_flip_probability=p) for d, p, _ in cfgs]
    batch_s = [stim_amd.Circuit(t).compile_detector_sampler(seed=1000 + k) for k, t in enumerate(texts)]
    plain_s = [stim_amd.Circuit(t).compile_detector_sampler(seed=1000 + k) for k, t in enumerate(texts)]
    for s in batch_s + plain_s:
        s.engine_info()
    shots = [c[2] for c in cfgs]
    batch = BatchSampler(list(zip(batch_s, shots)))
    assert batch.info()['kernel_jobs'] == 64
    outs = batch.run(use_graph=True)
    flags = N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES | N.MFS_DEVICE_OUTPUT
    L = N.lib()
    for k, ((bd, bo), s, n) in enumerate(zip(outs, plain_s, shots)):
        d = torch.empty_like(bd)
        o = torch.empty_like(bo)
        N.check(L.mfs_sample(s._h, n, flags, d.data_ptr(), d.stride(0), o.data_ptr(), o.stride(0)))
        assert torch.equal(d, bd) and torch.equal(o, bo), f'job {k}: d={cfgs[k][0]} shots={n}'
        del d, o
    # Rates of the small-distance jobs with enough shots against the oracle's exact rates.
    checked = 0
    for k, ((bd, bo), (d, p, n)) in enumerate(zip(outs, cfgs)):
        if d > 5 or n < 200000:
            continue
        exact = fo.exact_detector_rates(texts[k])
        D = len(exact) - 1
        counts = np.concatenate([_device_column_counts(bd, D), _device_column_counts(bo, 1)])
        z = _z(counts, n, exact)
        assert np.max(np.abs(z)) < _zcrit(len(z)), (k, d, p, n)
        checked += 1
    assert checked >= 3
    torch.cuda.synchronize()


# ---------------------------------------------------------------- serving pool of 200 circuits

def test_serving_pool_mix_equals_alone_and_exact_rates():
    import math
    import random
    import stim_amd
    from stim_amd import serve
    pool = serve.serving_pool(2026)
    assert len(pool) == 200
    samplers = [stim_amd.Circuit(text).compile_detector_sampler(seed=10_000 + k) for k, (_, text) in enumerate(pool)]
    for s in samplers:
        s.engine_info()
    rng = random.Random(7)
    reqs = []
    for i in range(80):
        c = rng.randrange(len(pool))
        reqs.append((c, int(math.exp(rng.uniform(math.log(1e3), math.log(2e5)))), 900 + i))
    # Two large requests on small circuits for the rate check.
    small = [k for k, (name, _) in enumerate(pool) if name.startswith(('surface d=3 ', 'repetition d=3 '))]
    assert small
    for j, c in enumerate(small[:2]):
        reqs.append((c, 1 << 20, 5000 + j))
    srv = serve.Server(budget_bytes=200e9, step_cost_bytes=2e8)
    ids = []
    for i, (c, n, seed) in enumerate(reqs):
        ids.append(srv.submit(samplers[c], n, seed))
        if i % 10 == 9:
            srv.step()   # requests keep arriving while others run
    while not all(srv.done(r) for r in ids):
        assert srv.step()
    mixed = {r: srv.result(r) for r in ids}
    assert srv.stats()['steps'] > 1
    # A sample of requests served alone (fresh server) gives the same rows.
    for i in list(range(0, 80, 9)) + [80, 81][:len(small[:2])]:
        c, n, seed = reqs[i]
        alone = serve.Server(budget_bytes=200e9)
        rid = alone.submit(samplers[c], n, seed)
        while not alone.done(rid):
            alone.step()
        d, o = alone.result(rid)
        np.testing.assert_array_equal(d, mixed[ids[i]][0], err_msg=f'request {i} ({pool[c][0]})')
        np.testing.assert_array_equal(o, mixed[ids[i]][1])
    for j, c in enumerate(small[:2]):
        i = 80 + j
        d, o = mixed[ids[i]]
        exact = fo.exact_detector_rates(pool[c][1])
        D = len(exact) - 1
        bits = np.unpackbits(d, axis=1, bitorder='little')[:, :D]
        counts = np.concatenate([bits.sum(axis=0), [int((o[:, 0] & 1).sum())]])
        z = _z(counts, reqs[i][1], exact)
        assert np.max(np.abs(z)) < _zcrit(len(z)), pool[c][0]
