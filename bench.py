#!/usr/bin/env python3
"""Benchmarks of the MI355X detection-event sampler. Each run prints one JSON line.

Default workload (the headline): detector shots/s of the d=25, 25-round rotated surface-code memory
circuit at p=1e-3 on one MI355X. One step = one sample of 2^20 shots (bit-packed detection events +
separate observables) into HBM-resident output rows.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--shots S] [--no-cpu] [--no-serve]

The default line also carries the metric's second half: a bounded replay of the serving workload (the
200-circuit pool, Poisson arrivals, --serve-requests requests at --qps) with served p50/p99 latency.
    python bench.py --workload batch64      # 64 heterogeneous surface codes: one persistent kernel vs 64 calls
    python bench.py --workload serve        # request loop with continuous batching, p50/p99 latency
    python bench.py --workload clifford     # 10,000-qubit, 1,000-layer random-Clifford circuit, 2^16 shots
    python bench.py --workload dem          # detector-error-model sampler, the d=25 r=25 model, 2^20 shots
    python bench.py --workload measure      # measurement sampler (compile_sampler), d=25 r=25, 2^20 shots
    python bench.py --workload m2d          # measurement -> detection event conversion, d=25 r=25, 2^20 shots
    python bench.py --workload sinter       # the drop-in call shape: sample(2^20, bit_packed, separate_observables) into numpy
"""

import argparse
import json
import math
import os
import random
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CIRCUIT = os.path.join(ROOT, 'tests', 'golden', 'circuits', 'surface_d25_r25.stim')
REF_STIM = os.path.join(ROOT, 'oracle', '_ref', 'stim')
TRAFFIC = os.path.join(ROOT, 'profiles', '{kernel}_d25_r25_2p20_traffic.md')
DEM_D25 = os.path.join(ROOT, 'tests', 'golden', 'dems', 'surface_d25_r25_folded.dem.gz')


# One reference process per host core, at most 16: a one-GPU lease's share of the host is 16 CPUs,
# even where the affinity mask shows more. Every baseline records both counts.
CPU_SHARE = 16


def _host_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cores():
    return max(1, min(_host_cpus(), CPU_SHARE))


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def _baseline(value, cores, sample):
    """A cpu_baseline record: the reference CLI, `cores` processes at once."""
    return {'value': value, 'unit': 'shots/s', 'cores': cores, 'kind': 'reference', 'sample': sample,
            'affinity_cpus': _host_cpus(), 'os_cpu_count': os.cpu_count(), 'cpu_model': _cpu_model(),
            'processes_cap': CPU_SHARE}


def cpu_baseline(circuit_path):
    """The reference CLI `stim detect` (built from the reference's sources into oracle/_ref), one process
    per core, on a bounded sample; the numpy oracle port when the CLI is absent."""
    cores = _cores()
    if os.path.exists(REF_STIM) and os.access(REF_STIM, os.X_OK):
        shots = 1 << 19
        procs = []
        t0 = time.perf_counter()
        for k in range(cores):
            procs.append(subprocess.Popen([REF_STIM, 'detect', '--shots', str(shots), '--in', circuit_path,
                                           '--out_format', 'b8', '--out', '/dev/null', '--obs_out', '/dev/null',
                                           '--obs_out_format', 'b8', '--seed', str(k)],
                                          stdout=subprocess.DEVNULL, stderr=subprocess.PIPE))
        ok = all(p.wait() == 0 for p in procs)
        wall = time.perf_counter() - t0
        if ok:
            return _baseline(cores * shots / wall, cores, f'stim detect b8 to /dev/null, {cores} processes x {shots} shots '
                              f'(d=25 r=25 p=1e-3), wall {wall:.2f}s')
    from oracle import frame_oracle as fo
    import numpy as np
    text = open(circuit_path).read()
    shots = 64
    t0 = time.perf_counter()
    fo.FrameOracle(text).run(shots, rng=np.random.default_rng(0))
    wall = time.perf_counter() - t0
    return {'value': shots / wall, 'unit': 'shots/s', 'cores': 1, 'kind': 'port',
            'sample': f'numpy oracle, 1 process x {shots} shots, wall {wall:.2f}s'}


def measured_traffic(kernel):
    """HBM bytes per launch of the headline kernel, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of that kernel (scaled by the 1 GiB copy calibration recorded there)."""
    try:
        text = open(TRAFFIC.format(kernel=kernel)).read()
    except OSError:
        return None
    m = re.search(r'^traffic_bytes_per_launch:\s*(\d+)', text, re.M)
    return int(m.group(1)) if m else None


def device_buffers(torch, shots, D, O):
    drow, orow = (D + 7) // 8, (O + 7) // 8
    dets = torch.empty((shots, max(drow, 1)), dtype=torch.uint8, device='cuda:0')
    obs = torch.empty((shots, max(orow, 1)), dtype=torch.uint8, device='cuda:0')
    return dets, obs, drow, orow


def headline(args):
    import torch
    import stim_amd
    from stim_amd import _native as N

    text = open(CIRCUIT).read()
    circuit = stim_amd.Circuit(text)
    sampler = circuit.compile_detector_sampler(seed=2026)
    if args.tile_words:
        sampler.set_tile_words(args.tile_words)
    D, O = sampler.num_detectors, sampler.num_observables
    S = args.shots
    dets, obs, drow, orow = device_buffers(torch, S, D, O)
    flags = N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES | N.MFS_DEVICE_OUTPUT
    if args.calibrate_copy:
        # A device-to-device copy of exactly 1 GiB: the reference dispatch for FETCH_SIZE / WRITE_SIZE.
        src = torch.ones(1 << 30, dtype=torch.uint8, device='cuda:0')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    L = N.lib()

    def step():
        N.check(L.mfs_sample(sampler._h, S, flags, dets.data_ptr(), dets.stride(0), obs.data_ptr(), obs.stride(0)))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = args.steps * S / wall

    # Kernel-only duration from HIP events on the sampler's own stream.
    kernel_ms = sampler.bench_device(S, warmup=5, iters=max(10, min(args.steps, 50)))
    engine = sampler.engine_info()
    other = 'frame' if engine['engine'] == 'symptom' else 'symptom'
    engines = {engine['engine']: {'shots_per_s': value, 'kernel_ms': kernel_ms}}
    try:
        alt = circuit.compile_detector_sampler(seed=2027)
        alt.set_engine(other)
        alt_ms = alt.bench_device(S, warmup=1, iters=3)
        engines[other] = {'shots_per_s': S / (alt_ms * 1e-3), 'kernel_ms': alt_ms}
    except ValueError:
        pass
    if engine['engine'] == 'symptom':
        # The generic scan kernel (basis-column records) on the same circuit, for comparison.
        gen_s = circuit.compile_detector_sampler(seed=2028)
        gen_s.set_engine('symptom')
        gen_s.set_scan_variant('generic')
        gen_ms = gen_s.bench_device(S, warmup=2, iters=10)
        engines['symptom_generic_scan'] = {'shots_per_s': S / (gen_ms * 1e-3), 'kernel_ms': gen_ms}
    b_alg = S * (drow + orow)
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    # The specialised scan kernel runs where the outcome table applies (outcome records built).
    kname = ('sym_scan_fast_kernel' if engine.get('outcome_records') else 'sym_scan_kernel') \
        if engine['engine'] == 'symptom' else 'frame_sim_kernel'
    traffic = measured_traffic(kname) if S == 1 << 20 and engine['engine'] == 'symptom' else None
    cpu = None if args.no_cpu else cpu_baseline(CIRCUIT)
    if engine['engine'] == 'symptom':
        tile = {'tile_shots': engine['tile_shots'], 'lds_bytes_per_tile': engine['lds_bytes_per_tile']}
    else:
        info = sampler.info()
        tile = {'tile_shots': 32 * info['tile_words'], 'lds_bytes_per_tile': info['lds_bytes']}
    return {
        'metric': 'detector shots/s @ d=25 r=25 p=1e-3 (HBM GB/s vs peak); served p50/p99 ms',
        'value': value,
        'unit': 'shots/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u32',
        'data': 'synthetic: rotated surface code memory-Z d=25 r=25 p=1e-3 (stim gen text), seeded Philox noise',
        'config': dict({'workload': 'surface_code:rotated_memory_z d=25 rounds=25 p=1e-3, 2^20 shots, '
                                    'bit_packed dets + separate observables, outputs resident in HBM',
                        'shots_per_step': S, 'detectors': D, 'observables': O, 'parallelism': 'single GPU'}, **tile),
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic,
                     'kernel': kname, 'kernel_ms': kernel_ms, 'algorithmic_bytes_per_launch': b_alg},
        'engine': engine,
        'engines': engines,
        'cpu_baseline': cpu,
    }


def add_serving(line, args):
    """Bounded serving replay (the 200-circuit pool) attached to the headline line."""
    from stim_amd import serve
    sv = serve.bench(qps=args.qps, requests=args.serve_requests)
    line['served_p50_ms'] = sv['latency_ms']['p50']
    line['served_p99_ms'] = sv['latency_ms']['p99']
    line['served_shots_per_s'] = sv['value']
    line['served'] = {k: sv[k] for k in ('latency_ms', 'qps_achieved', 'output_GBps', 'roofline', 'gpu_busy_frac',
                                          'steps', 'wall_s', 'compile_and_build_s', 'peak_bytes_in_use')}
    line['served']['qps_offered'] = args.qps
    line['served']['requests'] = args.serve_requests
    line['served']['data'] = sv['data']
    return line


def batch64_configs(seed=2026, n=64):
    """64 rotated memory-Z surface codes: d uniform over {3, 5, ..., 21}, rounds = d, p log-uniform over
    [1e-4, 5e-3], shots log-uniform over [1e4, 1e7]; the draw is seeded."""
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        d = rng.choice(range(3, 22, 2))
        p = math.exp(rng.uniform(math.log(1e-4), math.log(5e-3)))
        shots = int(math.exp(rng.uniform(math.log(1e4), math.log(1e7))))
        out.append((d, p, shots))
    return out


def batch64(args):
    import torch
    import stim_amd
    from stim_amd import _native as N
    from stim_amd import gen
    from stim_amd.batch import BatchSampler

    cfgs = batch64_configs()
    texts = [gen.generate('surface_code:rotated_memory_z', distance=d, rounds=d, after_clifford_depolarization=p,
                          after_reset_flip_probability=p, before_measure_flip_probability=p) for d, p, _ in cfgs]
    t0 = time.perf_counter()
    batch_s = [stim_amd.Circuit(t).compile_detector_sampler(seed=1000 + k) for k, t in enumerate(texts)]
    plain_s = [stim_amd.Circuit(t).compile_detector_sampler(seed=1000 + k) for k, t in enumerate(texts)]
    tune_s = [stim_amd.Circuit(t).compile_detector_sampler(seed=5000 + k) for k, t in enumerate(texts)]
    for s in batch_s + plain_s + tune_s:
        s.engine_info()      # builds the symptom tables
    compile_s = time.perf_counter() - t0
    shots = [c[2] for c in cfgs]
    total_shots = sum(shots)
    total_bytes = sum(n * ((s.num_detectors + 7) // 8 + (s.num_observables + 7) // 8) for s, n in zip(batch_s, shots))

    # Autotuner on its own samplers (so the compared samplers start at the same stream position).
    tuner = BatchSampler(list(zip(tune_s, shots)))
    kb, tune_ms = tuner.autotune(candidates=(12, 16, 20, 24, 32), iters=1)
    del tuner
    for s in batch_s + plain_s:
        s.set_symptom_tile_kb(kb)
    batch = BatchSampler(list(zip(batch_s, shots)))
    batch.set_tile_kb(kb)
    info = batch.info()

    # Results: the batch (graph replay) against 64 plain calls, bit for bit, on the device.
    outs = batch.run(use_graph=True)
    flags = N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES | N.MFS_DEVICE_OUTPUT
    L = N.lib()
    mismatches = 0
    for (bd, bo), s, n in zip(outs, plain_s, shots):
        d = torch.empty_like(bd)
        o = torch.empty_like(bo)
        N.check(L.mfs_sample(s._h, n, flags, d.data_ptr(), d.stride(0), o.data_ptr(), o.stride(0)))
        mismatches += int(not (torch.equal(d, bd) and torch.equal(o, bo)))
        del d, o
    torch.cuda.synchronize()

    # Timing: graph-replayed persistent kernel vs 64 separate plain calls (each call synchronous, as the
    # reference's sample() is).
    dev_ms, wall_ms = batch.time(use_graph=True, warmup=1, iters=args.steps)
    bufs = []
    for s, n in zip(plain_s, shots):
        d, o, _, _ = device_buffers(torch, n, s.num_detectors, s.num_observables)
        bufs.append((d, o))
    for (d, o), s, n in zip(bufs, plain_s, shots):
        N.check(L.mfs_sample(s._h, n, flags, d.data_ptr(), d.stride(0), o.data_ptr(), o.stride(0)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for (d, o), s, n in zip(bufs, plain_s, shots):
            N.check(L.mfs_sample(s._h, n, flags, d.data_ptr(), d.stride(0), o.data_ptr(), o.stride(0)))
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) / args.steps * 1e3
    return {
        'metric': 'batched detector shots/s: 64 heterogeneous surface codes in one persistent kernel',
        'value': total_shots / (wall_ms * 1e-3),
        'unit': 'shots/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': 1,
        'ms_per_step': wall_ms,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u32',
        'data': 'synthetic: 64 rotated memory-Z surface codes, d in {3..21}, r=d, p log-uniform [1e-4,5e-3], '
                'shots log-uniform [1e4,1e7] (seeded draw 2026)',
        'config': {'workload': 'batch64', 'circuits': len(cfgs), 'total_shots': total_shots,
                   'output_bytes': total_bytes, 'autotuned_tile_kb': kb, 'autotune_best_ms': tune_ms,
                   'parallelism': 'single GPU', **info},
        'batch_device_ms': dev_ms,
        'batch_wall_ms': wall_ms,
        'plain_64_calls_wall_ms': plain_ms,
        'speedup_vs_plain': plain_ms / wall_ms,
        'achieved_GBps': total_bytes / (dev_ms * 1e-3) / 1e9,
        'roofline': {'bound': 'hbm', 'achieved': total_bytes / (dev_ms * 1e-3) / 1e9, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': total_bytes / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 'traffic': None,
                     'kernel': 'sym_batch_kernel', 'kernel_ms': dev_ms, 'algorithmic_bytes_per_launch': total_bytes},
        'bit_exact_vs_plain': mismatches == 0,
        'mismatched_jobs': mismatches,
        'compile_and_build_s': compile_s,
        'cpu_baseline': None if args.no_cpu else jobs_cpu_baseline(texts, shots, 256, 'batch64: 64 surface codes'),
    }


def jobs_cpu_baseline(texts, shots, scale, what):
    """The reference CLI `stim detect` over a job mix (circuit text, shots), each job's shots divided by
    `scale` (at least 1,024), run by one worker per core pulling jobs in turn; shots/s = sampled shots /
    wall time."""
    import tempfile
    import concurrent.futures as cf
    if not (os.path.exists(REF_STIM) and os.access(REF_STIM, os.X_OK)):
        return None
    cores = _cores()
    with tempfile.TemporaryDirectory() as td:
        paths = {}
        jobs = []
        for t, n in zip(texts, shots):
            if t not in paths:
                paths[t] = os.path.join(td, f'c{len(paths)}.stim')
                with open(paths[t], 'w') as f:
                    f.write(t)
            jobs.append((paths[t], max(1024, int(n) // scale)))

        def one(job):
            path, n = job
            r = subprocess.run([REF_STIM, 'detect', '--shots', str(n), '--in', path, '--out_format', 'b8',
                                '--out', '/dev/null', '--obs_out', '/dev/null', '--obs_out_format', 'b8'],
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, stdin=subprocess.DEVNULL)
            return r.returncode == 0
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(cores) as ex:
            ok = all(ex.map(one, jobs))
        wall = time.perf_counter() - t0
    if not ok:
        return None
    total = sum(n for _, n in jobs)
    return _baseline(total / wall, cores, f'stim detect b8 to /dev/null over the {what} job mix, each job at '
                     f'1/{scale} of its shots (at least 1,024): {len(jobs)} jobs, {total} shots, {cores} workers, '
                     f'wall {wall:.2f}s')


def clifford(args):
    """Random-Clifford stress circuit: 10,000 qubits, 1,000 layers, 2^16 shots (frame engine)."""
    import torch
    import stim_amd
    from stim_amd import _native as N
    from stim_amd import gen
    t0 = time.perf_counter()
    text = gen.random_clifford(n_qubits=10000, layers=1000, seed=2026)
    sampler = stim_amd.Circuit(text).compile_detector_sampler(seed=7)
    prep_s = time.perf_counter() - t0
    S = 1 << 16
    D, O = sampler.num_detectors, sampler.num_observables
    dets, obs, drow, orow = device_buffers(torch, S, D, O)
    flags = N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES | N.MFS_DEVICE_OUTPUT
    L = N.lib()
    steps = max(1, min(args.steps, 3))
    N.check(L.mfs_sample(sampler._h, S, flags, dets.data_ptr(), dets.stride(0), obs.data_ptr(), obs.stride(0)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        N.check(L.mfs_sample(sampler._h, S, flags, dets.data_ptr(), dets.stride(0), obs.data_ptr(), obs.stride(0)))
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    kernel_ms = sampler.bench_device(S, warmup=1, iters=steps)
    info = sampler.info()
    return {
        'metric': 'detector shots/s, random-Clifford stress circuit (10,000 qubits x 1,000 layers)',
        'value': S / wall,
        'unit': 'shots/s',
        'n_gpus': 1,
        'steps': steps,
        'warmup': 1,
        'ms_per_step': wall * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u32',
        'data': 'synthetic: seeded random-Clifford circuit (H/S/S_DAG/SQRT_X, CX/CZ/SWAP/ISWAP on a random matching, '
                'PAULI_CHANNEL_1(3e-5 x3), DEPOLARIZE2(1e-4), MPP pairs of 500 weight-4 products every 10 layers, '
                'M/MR on 10% with CX rec feedback), seed 2026',
        'config': {'workload': 'clifford', 'qubits': sampler.num_qubits, 'detectors': D, 'observables': O,
                   'measurements': sampler.num_measurements, 'shots_per_step': S, 'engine': 'frame',
                   'tile_words': info['tile_words'], 'lds_bytes_per_tile': info['lds_bytes'],
                   'device_ops': info['ops'], 'parallelism': 'single GPU'},
        'output_GBps': S * (drow + orow) / wall / 1e9,
        'roofline': {'bound': 'hbm', 'achieved': S * (drow + orow) / (kernel_ms * 1e-3) / 1e9, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': S * (drow + orow) / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     'traffic': None, 'kernel': 'frame_sim_kernel<1>', 'kernel_ms': kernel_ms,
                     'algorithmic_bytes_per_launch': S * (drow + orow)},
        'generate_and_compile_s': prep_s,
        'cpu_baseline': None if args.no_cpu else clifford_cpu_baseline(text),
    }


def clifford_cpu_baseline(text):
    """The reference CLI `stim detect` on the same 10,000-qubit circuit, one process per core, 1,024
    shots each (one batch of the reference's 256-bit SIMD words)."""
    import tempfile
    if not (os.path.exists(REF_STIM) and os.access(REF_STIM, os.X_OK)):
        return None
    cores, shots = _cores(), 1024
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, 'clifford.stim')
        with open(path, 'w') as f:
            f.write(text)
        ok, wall = _run_parallel([[REF_STIM, 'detect', '--shots', str(shots), '--in', path, '--out_format', 'b8',
                                   '--out', '/dev/null', '--obs_out', '/dev/null', '--obs_out_format', 'b8',
                                   '--seed', str(k)] for k in range(cores)])
    if not ok:
        return None
    return _baseline(cores * shots / wall, cores, f'stim detect b8 to /dev/null, {cores} processes x {shots} shots '
                     f'(random Clifford, 10,000 qubits x 1,000 layers, includes parsing the circuit), '
                     f'wall {wall:.2f}s')


def dem_cpu_baseline(text):
    """The reference CLI `stim sample_dem`, one process per core, on a bounded sample of the same model."""
    import tempfile
    cores = _cores()
    if not (os.path.exists(REF_STIM) and os.access(REF_STIM, os.X_OK)):
        return None
    shots = 1 << 18
    with tempfile.NamedTemporaryFile('w', suffix='.dem', delete=False) as f:
        f.write(text)
        path = f.name
    try:
        procs = []
        t0 = time.perf_counter()
        for k in range(cores):
            procs.append(subprocess.Popen([REF_STIM, 'sample_dem', '--shots', str(shots), '--in', path,
                                           '--out_format', 'b8', '--out', '/dev/null', '--obs_out', '/dev/null',
                                           '--obs_out_format', 'b8', '--seed', str(k)],
                                          stdout=subprocess.DEVNULL, stderr=subprocess.PIPE))
        ok = all(p.wait() == 0 for p in procs)
        wall = time.perf_counter() - t0
    finally:
        os.unlink(path)
    if not ok:
        return None
    return _baseline(cores * shots / wall, cores, f'stim sample_dem b8 to /dev/null, {cores} processes x {shots} shots '
                      f'(d=25 r=25 p=1e-3 model), wall {wall:.2f}s')


def dem(args):
    """Detector-error-model sampler on the d=25, 25-round surface code's model with folded loops
    (356,321 mechanisms once unrolled)."""
    import gzip
    import torch
    import stim_amd
    from stim_amd import _native as N
    with gzip.open(DEM_D25, 'rt') as f:
        text = f.read()
    t0 = time.perf_counter()
    sampler = stim_amd.DetectorErrorModel(text).compile_sampler(seed=2026)
    compile_s = time.perf_counter() - t0
    D, O = sampler.num_detectors, sampler.num_observables
    S = args.shots
    dets, obs, drow, orow = device_buffers(torch, S, D, O)
    flags = N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES | N.MFS_DEVICE_OUTPUT
    L = N.lib()

    def step():
        N.check(L.mfs_sample(sampler._h, S, flags, dets.data_ptr(), dets.stride(0), obs.data_ptr(), obs.stride(0)))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = sampler.bench_device(S, warmup=5, iters=max(10, min(args.steps, 50)))
    engine = sampler.engine_info()
    b_alg = S * (drow + orow)
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    return {
        'metric': 'DEM sampler detector shots/s @ d=25 r=25 p=1e-3 model',
        'value': args.steps * S / wall,
        'unit': 'shots/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'u32',
        'data': 'the reference analyze_errors model (folded loops) of the d=25 r=25 p=1e-3 rotated memory-Z '
                'circuit, seeded Philox draws',
        'config': {'workload': 'dem_d25_r25', 'mechanisms': sampler.num_errors, 'detectors': D, 'observables': O,
                   'shots_per_step': S, 'tile_shots': engine['tile_shots'],
                   'lds_bytes_per_tile': engine['lds_bytes_per_tile'], 'probability_buckets': engine['classes'],
                   'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': None, 'kernel': 'sym_scan_waves_kernel',
                     'kernel_ms': kernel_ms, 'algorithmic_bytes_per_launch': b_alg},
        'compile_s': compile_s,
        'cpu_baseline': None if args.no_cpu else dem_cpu_baseline(text),
    }




def _run_parallel(cmds):
    procs = [subprocess.Popen(c, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, stdin=subprocess.DEVNULL)
             for c in cmds]
    t0 = time.perf_counter()
    ok = all(p.wait() == 0 for p in procs)
    return ok, time.perf_counter() - t0


def measure_cpu_baseline():
    """The reference CLI `stim sample` (measurement results), one process per core, bounded sample."""
    if not (os.path.exists(REF_STIM) and os.access(REF_STIM, os.X_OK)):
        return None
    cores, shots = _cores(), 1 << 18
    ok, wall = _run_parallel([[REF_STIM, 'sample', '--shots', str(shots), '--in', CIRCUIT, '--out_format', 'b8',
                               '--out', '/dev/null', '--seed', str(k)] for k in range(cores)])
    if not ok:
        return None
    return _baseline(cores * shots / wall, cores, f'stim sample b8 to /dev/null, {cores} processes x {shots} shots (d=25 r=25 p=1e-3), '
                      f'wall {wall:.2f}s')


def m2d_cpu_baseline():
    """The reference CLI `stim m2d` on a file of 2^19 sampled d=25 measurement shots, one process per core."""
    import tempfile
    if not (os.path.exists(REF_STIM) and os.access(REF_STIM, os.X_OK)):
        return None
    cores, shots = _cores(), 1 << 19
    with tempfile.TemporaryDirectory() as td:
        meas = os.path.join(td, 'm.b8')
        ok, _ = _run_parallel([[REF_STIM, 'sample', '--shots', str(shots), '--in', CIRCUIT, '--out_format', 'b8',
                                '--out', meas, '--seed', '1']])
        if not ok:
            return None
        ok, wall = _run_parallel([[REF_STIM, 'm2d', '--circuit', CIRCUIT, '--in', meas, '--in_format', 'b8',
                                   '--out', '/dev/null', '--out_format', 'b8', '--append_observables']
                                  for _ in range(cores)])
    if not ok:
        return None
    return _baseline(cores * shots / wall, cores, f'stim m2d b8 -> b8 (observables appended) to /dev/null, {cores} processes x {shots} shots '
                      f'(d=25 r=25 measurement file), wall {wall:.2f}s')


def _timed(step, args):
    import torch
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def measure(args):
    """Measurement sampler (every measurement result) on the d=25, 25-round circuit: the expansion path
    (detection events on the symptom engine, expanded into records) and, for comparison, the frame engine."""
    import torch
    import stim_amd
    from stim_amd import _native as N
    circuit = stim_amd.Circuit(open(CIRCUIT).read())
    t0 = time.perf_counter()
    sampler = circuit.compile_sampler(seed=2026)
    compile_s = time.perf_counter() - t0
    if args.tile_words:
        sampler.set_tile_words(args.tile_words)
    M, S = sampler.num_measurements, args.shots
    row = (M + 7) // 8
    out = torch.empty((S, row), dtype=torch.uint8, device='cuda:0')
    flags = N.MFS_BIT_PACKED | N.MFS_DEVICE_OUTPUT
    L = N.lib()

    def step():
        N.check(L.mfs_sample(sampler._h, S, flags, out.data_ptr(), row, None, 0))

    wall = _timed(step, args)
    iters = max(5, min(args.steps, 20))
    kernel_ms, last_ms = sampler.bench_device(S, warmup=3, iters=iters, split=True)
    info = sampler.info()
    engine = sampler.engine_info()
    engines = {engine['engine']: {'shots_per_s': args.steps * S / wall, 'kernel_ms': kernel_ms}}
    if engine['engine'] == 'expansion':
        frame = circuit.compile_sampler(seed=2027)
        frame.set_engine('frame')
        fms = frame.bench_device(S, warmup=2, iters=max(3, min(args.steps, 5)))
        engines['frame'] = {'shots_per_s': S / (fms * 1e-3), 'kernel_ms': fms}
        del frame
        # The expansion kernel reads the padded detection-event rows and writes the records.
        ypitch = ((circuit.num_detectors + circuit.num_observables + 7) // 8 + 15) // 16 * 16
        b_alg = S * (ypitch + row)
        achieved = b_alg / (last_ms * 1e-3) / 1e9
        roof = {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': achieved / HBM_PEAK_GBS, 'traffic': None, 'kernel': 'expand_kernel', 'kernel_ms': last_ms,
                'algorithmic_bytes_per_launch': b_alg,
                'symptom_kernel_ms': kernel_ms - last_ms}
        cfg = {'engine': 'expansion', 'symptom_columns': engine['columns'], 'free_results': engine['free_results'],
               'substitution_phases': engine['phases'], 'substitution_items': engine['items']}
    else:
        b_alg = S * row
        achieved = b_alg / (kernel_ms * 1e-3) / 1e9
        roof = {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': achieved / HBM_PEAK_GBS, 'traffic': None, 'kernel': 'frame_sim_kernel', 'kernel_ms': kernel_ms,
                'algorithmic_bytes_per_launch': b_alg}
        cfg = {'engine': 'frame', 'tile_shots': 32 * info['tile_words'], 'lds_bytes_per_tile': info['lds_bytes']}
    return {
        'metric': 'measurement shots/s @ d=25 r=25 p=1e-3 (all 16,225 results per shot)',
        'value': args.steps * S / wall, 'unit': 'shots/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'u32',
        'data': 'synthetic: rotated surface code memory-Z d=25 r=25 p=1e-3 (stim gen text), seeded Philox noise and '
                'uniform random results',
        'config': dict({'workload': 'measure_d25_r25', 'measurements': M, 'shots_per_step': S,
                        'parallelism': 'single GPU'}, **cfg),
        'roofline': roof,
        'engines': engines,
        'reference_sample_and_compile_s': compile_s,
        'cpu_baseline': None if args.no_cpu else measure_cpu_baseline(),
    }


def m2d(args):
    """Measurement -> detection event conversion of HBM-resident d=25 measurement rows."""
    import torch
    import stim_amd
    from stim_amd import _native as N
    circuit = stim_amd.Circuit(open(CIRCUIT).read())
    conv = circuit.compile_m2d_converter()
    M, D, O, S = conv.num_measurements, conv.num_detectors, conv.num_observables, args.shots
    in_stride = 16 * ((M + 127) // 128)
    out_row = (D + O + 7) // 8
    meas = torch.zeros((S, in_stride), dtype=torch.uint8, device='cuda:0')
    # Realistic input: sampled measurement rows written straight into the padded device rows.
    sampler = circuit.compile_sampler(seed=7)
    N.check(N.lib().mfs_sample(sampler._h, S, N.MFS_BIT_PACKED | N.MFS_DEVICE_OUTPUT, meas.data_ptr(), in_stride,
                               None, 0))
    torch.cuda.synchronize()
    out = torch.empty((S, out_row), dtype=torch.uint8, device='cuda:0')
    flags = N.MFS_BIT_PACKED | N.MFS_DEVICE_OUTPUT | N.MFS_APPEND_OBSERVABLES
    L = N.lib()

    def step():
        N.check(L.mfs_m2d_convert(conv._h, S, meas.data_ptr(), in_stride, flags, out.data_ptr(), out_row, None, 0))

    wall = _timed(step, args)
    info = conv.kernel_info()
    iters = max(10, min(args.steps, 50))
    kernels = {}
    for name in ('whole', 'windowed'):
        conv.set_kernel(name)
        kernels[name] = conv.bench_device(S, warmup=3, iters=iters)
    conv.set_kernel('auto')
    kernel_ms = conv.bench_device(S, warmup=3, iters=iters)
    # The default picks the whole-record kernel wherever a tile's record fits in LDS (160 KiB).
    kname = 'm2d_kernel' if info['lds_whole'] <= 160 * 1024 else 'm2d_win_kernel'
    b_alg = S * ((M + 7) // 8 + out_row)
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    return {
        'metric': 'converted shots/s @ d=25 r=25 (measurements -> detection events + observables)',
        'value': args.steps * S / wall, 'unit': 'shots/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'u32',
        'data': 'synthetic: measurement rows sampled on the GPU from the d=25 r=25 p=1e-3 circuit, resident in HBM',
        'config': {'workload': 'm2d_d25_r25', 'measurements': M, 'detectors': D, 'observables': O,
                   'shots_per_step': S, 'tile_shots': 32, 'ring_words': info['ring_words_appended'],
                   'lds_bytes_per_tile': info['lds_whole'] if kname == 'm2d_kernel' else info['lds_windowed'],
                   'parallelism': 'single GPU'},
        'kernel_ms_by_kernel': {'m2d_kernel (whole record in LDS)': kernels['whole'],
                                'm2d_win_kernel (ring of recent results)': kernels['windowed']},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': None, 'kernel': kname,
                     'kernel_ms': kernel_ms, 'algorithmic_bytes_per_launch': b_alg},
        'cpu_baseline': None if args.no_cpu else m2d_cpu_baseline(),
    }


def sinter(args):
    """The reference's own call shape as sinter makes it (src/stim/py/compiled_detector_sampler.pybind.cc:34-86;
    glue/sample/src/sinter/_decoding/_stim_then_decode_sampler.py:162-174): sample(2^20, bit_packed=True,
    separate_observables=True) returning numpy arrays in host memory, d=25 r=25 p=1e-3."""
    import numpy as np
    import stim_amd
    circuit = stim_amd.Circuit(open(CIRCUIT).read())
    sampler = circuit.compile_detector_sampler(seed=2026)
    S = args.shots
    D, O = sampler.num_detectors, sampler.num_observables
    dets = np.empty((S, (D + 7) // 8), dtype=np.uint8)
    obs = np.empty((S, (O + 7) // 8), dtype=np.uint8)
    steps = max(1, min(args.steps, 10))
    for _ in range(max(1, min(args.warmup, 3))):
        sampler.sample(S, bit_packed=True, separate_observables=True, dets_out=dets, obs_out=obs)
    t0 = time.perf_counter()
    for _ in range(steps):
        sampler.sample(S, bit_packed=True, separate_observables=True, dets_out=dets, obs_out=obs)
    wall = (time.perf_counter() - t0) / steps
    kernel_ms = sampler.bench_device(S, warmup=2, iters=10)
    out_bytes = S * (dets.shape[1] + obs.shape[1])
    return {
        'metric': 'detector shots/s into host numpy arrays, sinter call shape, d=25 r=25 p=1e-3',
        'value': S / wall, 'unit': 'shots/s', 'n_gpus': 1, 'steps': steps, 'warmup': 1,
        'ms_per_step': wall * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'u32',
        'data': 'synthetic: rotated surface code memory-Z d=25 r=25 p=1e-3, seeded Philox noise',
        'config': {'workload': 'sinter_sample_numpy', 'shots_per_call': S, 'bit_packed': True,
                   'separate_observables': True, 'output': 'preallocated numpy arrays (dets_out/obs_out), pageable',
                   'parallelism': 'single GPU'},
        'host_GBps': out_bytes / wall / 1e9,
        'pcie_peak_GBps': 64.0,
        'pcie_frac': out_bytes / wall / 1e9 / 64.0,
        'kernel_ms': kernel_ms,
        'kernel_share_of_call': kernel_ms / (wall * 1e3),
        'cpu_baseline': None if args.no_cpu else cpu_baseline(CIRCUIT),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--shots', type=int, default=1 << 20)
    ap.add_argument('--workload', choices=['surface_d25', 'batch64', 'serve', 'clifford', 'dem', 'measure', 'm2d',
                                           'sinter'], default='surface_d25')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--tile-words', type=int, default=0)
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='run a 1 GiB device-to-device copy first (reference for FETCH_SIZE/WRITE_SIZE counters)')
    ap.add_argument('--qps', type=float, default=300.0)
    ap.add_argument('--requests', type=int, default=3000)
    ap.add_argument('--serve-requests', type=int, default=1000,
                    help='requests of the serving replay attached to the default line')
    ap.add_argument('--no-serve', action='store_true', help='leave the serving replay out of the default line')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path (shots would shard across GPUs '
              'with no exchange). Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)
    if args.workload == 'surface_d25':
        line = headline(args)
        if not args.no_serve:
            line = add_serving(line, args)
    elif args.workload == 'batch64':
        args.steps = min(args.steps, 5)
        line = batch64(args)
    elif args.workload == 'clifford':
        line = clifford(args)
    elif args.workload == 'dem':
        line = dem(args)
    elif args.workload == 'measure':
        args.steps = min(args.steps, 10)
        args.warmup = min(args.warmup, 2)
        line = measure(args)
    elif args.workload == 'm2d':
        line = m2d(args)
    elif args.workload == 'sinter':
        line = sinter(args)
    else:
        from stim_amd import serve
        trace = {}
        line = serve.bench(qps=args.qps, requests=args.requests, trace=trace)
        if not args.no_cpu:
            texts = [trace['pool'][c][1] for c in trace['chosen'][:64]]
            line['cpu_baseline'] = jobs_cpu_baseline(texts, trace['shots'][:64], 64,
                                                     'serving trace (first 64 requests)')
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
