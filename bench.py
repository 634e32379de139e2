#!/usr/bin/env python3
"""Benchmark: detector shots/s of the d=25, 25-round rotated surface code at p=1e-3 on one MI355X.

One step = one GPU sample of 2^20 shots (bit-packed detection events + separate observables) into
HBM-resident output rows. Prints one JSON line (see the "metric"/"roofline"/"cpu_baseline" keys).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--shots S] [--no-cpu]
"""

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CIRCUIT = os.path.join(ROOT, 'tests', 'golden', 'circuits', 'surface_d25_r25.stim')
REF_STIM = os.path.join(ROOT, 'oracle', '_ref', 'stim')


def cpu_baseline(circuit_path, n_det, n_obs):
    """Reference CLI `stim detect` (built from the reference's sources into oracle/_ref), one process
    per core, on a bounded sample; falls back to the numpy oracle when the CLI is absent."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
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
            return {'value': cores * shots / wall, 'unit': 'shots/s', 'cores': cores, 'kind': 'reference',
                    'sample': f'stim detect b8 to /dev/null, {cores} processes x {shots} shots '
                              f'(d=25 r=25 p=1e-3), wall {wall:.2f}s'}
    from oracle import frame_oracle as fo
    text = open(circuit_path).read()
    shots = 64
    t0 = time.perf_counter()
    fo.FrameOracle(text).run(shots, rng=__import__('numpy').random.default_rng(0))
    wall = time.perf_counter() - t0
    return {'value': shots / wall, 'unit': 'shots/s', 'cores': 1, 'kind': 'port',
            'sample': f'numpy oracle, 1 process x {shots} shots, wall {wall:.2f}s'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--shots', type=int, default=1 << 20)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--tile-words', type=int, default=0)
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='run a 1 GiB device-to-device copy first (reference for FETCH_SIZE/WRITE_SIZE counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path (shots would shard '
              'across GPUs with no exchange; see DESIGN.md). Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

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
    drow, orow = (D + 7) // 8, (O + 7) // 8
    dev = torch.device('cuda:0')
    dets = torch.empty((S, drow), dtype=torch.uint8, device=dev)
    obs = torch.empty((S, max(orow, 1)), dtype=torch.uint8, device=dev)
    flags = N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES | N.MFS_DEVICE_OUTPUT
    if args.calibrate_copy:
        src = torch.ones(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    L = N.lib()

    def step():
        N.check(L.mfs_sample(sampler._h, S, flags, dets.data_ptr(), drow, obs.data_ptr(), max(orow, 1)))

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
    kernel_ms = sampler.bench_device(S, warmup=1, iters=max(3, min(args.steps, 10)))
    engine = sampler.engine_info()
    # The other engine, measured the same way, for comparison.
    other = 'frame' if engine['engine'] == 'symptom' else 'symptom'
    engines = {engine['engine']: {'shots_per_s': value, 'kernel_ms': kernel_ms}}
    try:
        alt = circuit.compile_detector_sampler(seed=2027)
        alt.set_engine(other)
        alt_ms = alt.bench_device(S, warmup=1, iters=3)
        engines[other] = {'shots_per_s': S / (alt_ms * 1e-3), 'kernel_ms': alt_ms}
    except ValueError:
        pass
    b_alg = S * (drow + orow)
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    info = sampler.info()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(CIRCUIT, D, O)

    line = {
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
        'dtype': 'u32 (GF(2) bit words)',
        'data': 'synthetic: rotated surface code memory-Z d=25 r=25 p=1e-3 (stim gen text), seeded Philox noise',
        'config': {'workload': 'surface_code:rotated_memory_z d=25 rounds=25 p=1e-3, 2^20 shots, '
                               'bit_packed dets + separate observables, outputs resident in HBM',
                   'shots_per_step': S, 'detectors': D, 'observables': O,
                   'tile_shots': 32 * info['tile_words'], 'lds_bytes_per_tile': info['lds_bytes'],
                   'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': None,
                     'kernel': 'sym_sample_kernel' if engine['engine'] == 'symptom' else 'frame_sim_kernel',
                     'kernel_ms': kernel_ms, 'algorithmic_bytes_per_launch': b_alg},
        'engine': engine,
        'engines': engines,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
