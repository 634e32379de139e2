"""Serving: resident compiled circuits, memory-budget admission and continuous batching.

No reference counterpart (the reference samples synchronously, one call at a time). A `Server` keeps
compiled samplers (their programs and symptom tables stay in HBM) and takes requests
(sampler, shots, seed). A request is admitted once its output rows fit the memory budget; every
step runs a fair share of every admitted request's shot tiles in one persistent-kernel launch, so
requests that arrive while others run join the next step. A request's result depends only on its
circuit, seed and shot count, not on what it was batched with.

    srv = Server(budget_bytes=200e9)
    rid = srv.submit(sampler, shots=100_000, seed=7)
    while not srv.done(rid): srv.step()
    dets, obs = srv.result(rid)        # torch-free: numpy copies; srv.release(rid) frees the rows
"""

import ctypes
import math
import random

import numpy as np

from . import _native as N


def _bind():
    return N.lib()


class Server:
    def __init__(self, budget_bytes=200e9, step_cost_bytes=2e9):
        L = _bind()
        h = ctypes.c_void_p()
        N.check(L.mfs_server_create(int(budget_bytes), float(step_cost_bytes), ctypes.byref(h)))
        self._h = h
        self._samplers = {}
        self._done = set()
        self.times = {}

    def __del__(self):
        if getattr(self, '_h', None):
            N.lib().mfs_server_destroy(self._h)
            self._h = None

    def submit(self, sampler, shots, seed):
        rid = ctypes.c_uint64()
        N.check(N.lib().mfs_server_submit(self._h, sampler._h, int(shots), int(seed), ctypes.byref(rid)))
        self._samplers[rid.value] = sampler
        return rid.value

    def step(self):
        """Runs one scheduling step; returns False when nothing was queued or running."""
        w = ctypes.c_int()
        N.check(N.lib().mfs_server_step(self._h, ctypes.byref(w)))
        ids = (ctypes.c_uint64 * 1024)()
        n = ctypes.c_uint32()
        while True:
            N.check(N.lib().mfs_server_poll(self._h, ids, 1024, ctypes.byref(n)))
            self._done.update(ids[i] for i in range(n.value))
            if n.value < 1024:
                break
        return bool(w.value)

    def done(self, rid):
        return rid in self._done

    def result(self, rid):
        """Copies a finished request's bit-packed rows to host: (dets, obs) uint8 arrays. The request's
        timestamps (submitted, admitted, finished; seconds) are kept in `self.times[rid]`."""
        dp, op = ctypes.c_void_p(), ctypes.c_void_p()
        ds, os_ = ctypes.c_size_t(), ctypes.c_size_t()
        shots = ctypes.c_uint64()
        times = (ctypes.c_double * 3)()
        N.check(N.lib().mfs_server_result(self._h, int(rid), ctypes.byref(dp), ctypes.byref(ds), ctypes.byref(op),
                                          ctypes.byref(os_), ctypes.byref(shots), times))
        self.times[rid] = tuple(times)
        s = self._samplers[rid]
        drow, orow = (s.num_detectors + 7) // 8, (s.num_observables + 7) // 8
        dets = np.empty((shots.value, drow), dtype=np.uint8)
        obs = np.empty((shots.value, orow), dtype=np.uint8)
        N.check(N.lib().mfs_server_copy_result(self._h, int(rid), dets.ctypes.data, max(drow, 1), obs.ctypes.data,
                                               max(orow, 1)))
        return dets, obs

    def release(self, rid):
        N.check(N.lib().mfs_server_release(self._h, int(rid)))
        self._done.discard(rid)
        self._samplers.pop(rid, None)

    def stats(self):
        out = (ctypes.c_uint64 * 8)()
        N.check(N.lib().mfs_server_stats(self._h, out))
        v = list(out)
        d = dict(zip(['steps', 'queued', 'active', 'bytes_in_use', 'peak_bytes_in_use', 'budget'], v[:6]))
        d['kernel_ms'] = v[6] / 1000.0
        d['bytes_written'] = v[7]
        return d

    def replay(self, samplers, shots, seeds, arrival_s):
        """Runs a request trace in real time (arrival times in seconds from now); each result is
        released as soon as it is complete. Returns (latency ms per request, wall seconds)."""
        n = len(samplers)
        S = (ctypes.c_void_p * n)(*[s._h.value for s in samplers])
        sh = (ctypes.c_uint64 * n)(*[int(x) for x in shots])
        se = (ctypes.c_uint64 * n)(*[int(x) for x in seeds])
        ar = (ctypes.c_double * n)(*[float(x) for x in arrival_s])
        lat = (ctypes.c_double * n)()
        wall = ctypes.c_double()
        N.check(N.lib().mfs_server_replay(self._h, n, S, sh, se, ar, lat, ctypes.byref(wall)))
        return np.array(list(lat)), wall.value


def serving_pool(seed=2026, n_surface=80, n_rep=80, n_clifford=40, clifford_layers=20):
    """The serving pool: rotated surface codes d in {3..25} (odd), repetition codes d in {3..51}, and
    random-Clifford circuits of 500..5000 qubits, each at p log-uniform over [1e-4, 1e-2]."""
    from . import gen
    rng = random.Random(seed)

    def logu(lo, hi):
        return math.exp(rng.uniform(math.log(lo), math.log(hi)))
    pool = []
    for _ in range(n_surface):
        d, p = rng.choice(range(3, 26, 2)), logu(1e-4, 1e-2)
        pool.append((f'surface d={d} p={p:.2e}', gen.generate(
            'surface_code:rotated_memory_z', distance=d, rounds=d, after_clifford_depolarization=p,
            after_reset_flip_probability=p, before_measure_flip_probability=p)))
    for _ in range(n_rep):
        d, p = rng.randint(3, 51), logu(1e-4, 1e-2)
        pool.append((f'repetition d={d} p={p:.2e}', gen.generate(
            'repetition_code:memory', distance=d, rounds=d, after_clifford_depolarization=p,
            after_reset_flip_probability=p, before_measure_flip_probability=p)))
    for k in range(n_clifford):
        q, p = rng.randrange(500, 5001, 2), logu(1e-4, 1e-2)
        pool.append((f'random-clifford q={q} p={p:.2e}', gen.random_clifford(
            n_qubits=q, layers=clifford_layers, seed=k, px=p / 3, py=p / 3, pz=p / 3, p2=p)))
    return pool


def bench(qps=40.0, requests=400, seed=2026, budget_bytes=200e9, trace=None):
    """Poisson request stream over the serving pool; returns the benchmark's JSON record. When `trace` is
    a dict it receives the pool and the requests (circuit index and shots of each)."""
    import time
    import stim_amd
    t0 = time.perf_counter()
    pool = serving_pool(seed)
    samplers = [stim_amd.Circuit(text).compile_detector_sampler(seed=10_000 + k) for k, (_, text) in enumerate(pool)]
    for s in samplers:
        s.engine_info()
    compile_s = time.perf_counter() - t0
    rng = random.Random(seed + 1)
    t, arrivals, chosen, shots = 0.0, [], [], []
    for _ in range(requests):
        t += rng.expovariate(qps)
        arrivals.append(t)
        chosen.append(rng.randrange(len(samplers)))
        shots.append(int(math.exp(rng.uniform(math.log(1e3), math.log(1e7)))))
    if trace is not None:
        trace.update(pool=pool, chosen=chosen, shots=shots, samplers=samplers)
    srv = Server(budget_bytes=budget_bytes)
    lat, wall = srv.replay([samplers[c] for c in chosen], shots, list(range(requests)), arrivals)
    total_shots = sum(shots)
    out_bytes = sum(n * ((samplers[c].num_detectors + 7) // 8 + (samplers[c].num_observables + 7) // 8)
                    for c, n in zip(chosen, shots))
    st = srv.stats()
    return {
        'metric': 'served detector shots/s with p50/p99 latency (Poisson arrivals, continuous batching)',
        'value': total_shots / wall,
        'unit': 'shots/s',
        'n_gpus': 1,
        'steps': st['steps'],
        'warmup': 0,
        'ms_per_step': wall / max(st['steps'], 1) * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u32',
        'data': f'synthetic: pool of {len(pool)} circuits (80 surface d 3..25, 80 repetition d 3..51, '
                f'40 random-Clifford 500..5000 qubits x 20 layers), p log-uniform [1e-4,1e-2]; '
                f'{requests} requests, shots log-uniform [1e3,1e7], seeded',
        'config': {'workload': 'serve', 'qps_offered': qps, 'requests': requests, 'budget_bytes': int(budget_bytes),
                   'parallelism': 'single GPU'},
        'latency_ms': {'p50': float(np.percentile(lat, 50)), 'p99': float(np.percentile(lat, 99)),
                       'mean': float(lat.mean()), 'max': float(lat.max())},
        'qps_achieved': requests / wall,
        'output_GBps': out_bytes / wall / 1e9,
        'roofline': {'bound': 'hbm', 'achieved': st['bytes_written'] / max(st['kernel_ms'], 1e-9) * 1e-6,
                     'peak': 8000.0, 'unit': 'GB/s',
                     'frac': st['bytes_written'] / max(st['kernel_ms'], 1e-9) * 1e-6 / 8000.0, 'traffic': None,
                     'kernel': 'sym_batch_kernel', 'kernel_ms': st['kernel_ms'] / max(st['steps'], 1),
                     'algorithmic_bytes_per_launch': st['bytes_written'] / max(st['steps'], 1)},
        'gpu_busy_frac': st['kernel_ms'] / (wall * 1e3),
        'peak_bytes_in_use': st['peak_bytes_in_use'],
        'compile_and_build_s': compile_s,
        'wall_s': wall,
    }
