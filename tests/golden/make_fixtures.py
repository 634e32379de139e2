#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ from the reference Stim CLI.

The CLI is the reference's own sampler, built from the sources under /root/reference by
oracle/Makefile into oracle/_ref/stim. This script runs only in the build container (it refuses
to run when /root/reference or the built CLI is absent). Everything it writes is data:
circuit text produced by `stim gen`, fault lists with the detection events `stim detect` printed
for them, per-detector rates derived from `stim analyze_errors`, seeded firing counts, and
`stim detect` output files in every result format.

    make -C oracle -j8 ref && python tests/golden/make_fixtures.py
"""

import gzip
import json
import os
import random
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import frame_oracle as fo  # noqa: E402
from oracle import stim_text as st  # noqa: E402
import zoo  # noqa: E402

STIM = os.path.join(ROOT, 'oracle', '_ref', 'stim')


def run(args, stdin=None):
    return subprocess.run([STIM] + args, input=stdin, capture_output=True, check=True, text=True).stdout


def gen(code, task, d, r, p, extra=()):
    args = ['gen', '--code', code, '--task', task, '--distance', str(d), '--rounds', str(r)]
    if p:
        args += ['--after_clifford_depolarization', str(p), '--after_reset_flip_probability', str(p),
                 '--before_measure_flip_probability', str(p)]
    args += list(extra)
    return run(args)


def fmt_args(a):
    return '(' + ', '.join(repr(float(x)) for x in a) + ')' if a else ''


PAULI_OF = {1: 'X', 2: 'Z', 3: 'Y'}


def materialize(flat, faults):
    """Flat circuit text with all noise stripped and the given faults made deterministic.

    A fault (k, j, code) becomes X/Y/Z_ERROR(1), E(1) on the pair, a p=1 noisy result on that one
    result, a p=1 link of a correlated chain, or a p=1 heralded Pauli channel outcome."""
    fk = {}
    for (k, j, c) in faults:
        fk.setdefault(k, {})
        assert j not in fk[k], 'one fault per site'
        fk[k][j] = c
    out = []
    for k, ins in enumerate(flat):
        n, T, a = ins.name, ins.targets, ins.args
        F = fk.get(k, {})
        tstr = lambda ts: ' '.join(repr(t) for t in ts)  # noqa: E731
        if n in ('X_ERROR', 'Y_ERROR', 'Z_ERROR', 'DEPOLARIZE1', 'PAULI_CHANNEL_1', 'I_ERROR'):
            for j, c in sorted(F.items()):
                out.append(f'{PAULI_OF[c & 3]}_ERROR(1) {T[j].value}')
        elif n in ('DEPOLARIZE2', 'PAULI_CHANNEL_2', 'II_ERROR'):
            for j, c in sorted(F.items()):
                terms = []
                if c & 3:
                    terms.append(f'{PAULI_OF[c & 3]}{T[2 * j].value}')
                if (c >> 2) & 3:
                    terms.append(f'{PAULI_OF[(c >> 2) & 3]}{T[2 * j + 1].value}')
                out.append('E(1) ' + ' '.join(terms))
        elif n in st.MEASURE_1Q or n == 'MPAD':
            for j, t in enumerate(T):
                out.append(f'{n}{"(1)" if j in F else ""} {t!r}')
        elif n in st.PAIR_MEASURE:
            for j in range(len(T) // 2):
                out.append(f'{n}{"(1)" if j in F else ""} {T[2 * j]!r} {T[2 * j + 1]!r}')
        elif n == 'MPP':
            for j, g in enumerate(st.product_groups(T)):
                out.append(f'MPP{"(1)" if j in F else ""} ' + '*'.join(repr(t) for t in g))
        elif n in ('E', 'ELSE_CORRELATED_ERROR'):
            out.append(f'{n}({1 if 0 in F else 0}) {tstr(T)}')
        elif n in ('HERALDED_ERASE', 'HERALDED_PAULI_CHANNEL_1'):
            for j, t in enumerate(T):
                if j in F:
                    c = F[j]
                    pv = [0, 0, 0, 0]
                    pv[{0: 0, 1: 1, 3: 2, 2: 3}[c & 3]] = 1
                    out.append(f'HERALDED_PAULI_CHANNEL_1({pv[0]}, {pv[1]}, {pv[2]}, {pv[3]}) {t!r}')
                else:
                    out.append(f'HERALDED_ERASE(0) {t!r}')
        else:
            out.append(f'{n}{fmt_args(a)} {tstr(T)}'.rstrip())
    return '\n'.join(out) + '\n'


def detect_01(text, shots=1):
    with tempfile.NamedTemporaryFile('w', suffix='.stim', delete=False) as f:
        f.write(text)
        path = f.name
    try:
        return run(['detect', '--shots', str(shots), '--in', path, '--out_format', '01', '--append_observables'])
    finally:
        os.unlink(path)


def random_faults(flat, rng, count):
    chans = fo.noise_sites(flat, allow_chains=True)
    picks = rng.sample(chans, k=min(count, len(chans)))
    return [(k, j, rng.choice(outs)[0]) for (k, j, outs) in picks]


def fault_fixture(name, text, n_sets, rng, max_faults=4):
    flat = st.parse_flat(text)
    stats = st.stats(flat)
    noiseless = materialize(flat, [])
    base = detect_01(noiseless, shots=256).split('\n')
    assert all(set(line) <= {'0'} for line in base if line), f'{name}: nondeterministic detectors'
    sets = []
    for i in range(n_sets):
        nf = 1 if i < n_sets // 2 else rng.randint(2, max_faults)
        faults = random_faults(flat, rng, nf)
        line = detect_01(materialize(flat, faults)).strip()
        sets.append({'faults': [list(f) for f in faults], 'expected': line})
    return {'name': name, 'circuit': text, 'stats': stats, 'sets': sets}


def dem_rates(text):
    """Per-detector/observable firing probability implied by the reference's detector error model."""
    with tempfile.NamedTemporaryFile('w', suffix='.stim', delete=False) as f:
        f.write(text)
        path = f.name
    try:
        dem = run(['analyze_errors', '--in', path, '--approximate_disjoint_errors'])
    finally:
        os.unlink(path)
    stats = st.stats(st.parse_flat(text))
    nd, no = stats['num_detectors'], stats['num_observables']
    import math
    logs = [0.0] * (nd + no)
    off = 0
    for line in dem.split('\n'):
        line = line.strip()
        if line.startswith('shift_detectors'):
            off += int(line.split()[-1])
        elif line.startswith('error('):
            p = float(line[6:line.index(')')])
            for tok in line[line.index(')') + 1:].split():
                if tok.startswith('D'):
                    logs[off + int(tok[1:])] += math.log(1 - 2 * p)
                elif tok.startswith('L'):
                    logs[nd + int(tok[1:])] += math.log(1 - 2 * p)
        elif line.startswith('repeat'):
            raise RuntimeError('unexpected folded DEM')
    return [(1 - math.exp(v)) / 2 for v in logs]


def seeded_counts(path, shots, seed):
    out = subprocess.run([STIM, 'detect', '--shots', str(shots), '--in', path, '--out_format', 'b8',
                          '--append_observables', '--seed', str(seed)], capture_output=True, check=True).stdout
    import numpy as np
    stats = st.stats(st.parse_flat(open(path).read()))
    n = stats['num_detectors'] + stats['num_observables']
    arr = np.frombuffer(out, dtype=np.uint8).reshape(shots, (n + 7) // 8)
    bits = np.unpackbits(arr, axis=1, bitorder='little')[:, :n]
    return bits.sum(axis=0).astype(int).tolist()


GEN_PARAM_FLAGS = ['after_clifford_depolarization', 'before_round_data_depolarization',
                   'before_measure_flip_probability', 'after_reset_flip_probability']


def write_gen_pins():
    """stim gen text for the families the build's own generator reproduces, including the edge cases
    of its concatenation rules (one round, two rounds, zero noise, unequal noise, six-digit printing)."""
    def uniform(p):
        return {'after_clifford_depolarization': p, 'before_measure_flip_probability': p,
                'after_reset_flip_probability': p} if p else {}
    cases = [('surface_code', 'rotated_memory_z', 3, 3, uniform(1e-3)),
             ('surface_code', 'rotated_memory_z', 7, 4, uniform(0.002)),
             ('surface_code', 'rotated_memory_z', 11, 11, uniform(1e-4)),
             ('surface_code', 'rotated_memory_x', 5, 5, uniform(0.003)),
             ('repetition_code', 'memory', 3, 3, uniform(1e-3)),
             ('repetition_code', 'memory', 11, 7, uniform(0.005)),
             ('surface_code', 'rotated_memory_z', 5, 1, uniform(1e-3)),
             ('surface_code', 'rotated_memory_z', 4, 2, uniform(0.00123456789)),
             ('surface_code', 'rotated_memory_z', 2, 3, {}),
             ('surface_code', 'rotated_memory_x', 3, 2, {'after_clifford_depolarization': 0.01,
                                                         'before_round_data_depolarization': 0.02,
                                                         'before_measure_flip_probability': 0.03,
                                                         'after_reset_flip_probability': 0.04}),
             ('surface_code', 'unrotated_memory_z', 3, 2, uniform(1e-3)),
             ('surface_code', 'unrotated_memory_x', 4, 3, {'before_round_data_depolarization': 0.001}),
             ('repetition_code', 'memory', 2, 1, {}),
             ('repetition_code', 'memory', 51, 51, uniform(0.0007777777))]
    pins = []
    for (code, task, d, r, params) in cases:
        args = ['gen', '--code', code, '--task', task, '--distance', str(d), '--rounds', str(r)]
        for k, v in params.items():
            args += ['--' + k, repr(v)]
        pins.append({'code': code, 'task': task, 'distance': d, 'rounds': r, 'params': params,
                     'text': run(args)})
    json.dump(pins, open(os.path.join(HERE, 'gen_pins.json'), 'w'), indent=0)


def write_dem_fixtures():
    """Detector error models of fixture circuits (`stim analyze_errors`, unfolded and with REPEAT
    blocks folded) and replay vectors: random error masks fed to `stim sample_dem --replay_err_in`."""
    ddir = os.path.join(HERE, 'dems')
    os.makedirs(ddir, exist_ok=True)
    rng = random.Random(20261203)
    cases = []
    for name in ['surface_d3_r3', 'surface_d5_r5', 'rep_d7_r5', 'color_d5_r3']:
        path = os.path.join(HERE, 'circuits', name + '.stim')
        for folded in (False, True):
            args = ['analyze_errors', '--in', path, '--approximate_disjoint_errors']
            if folded:
                args.append('--fold_loops')
            text = run(args)
            dname = name + ('_folded' if folded else '')
            with open(os.path.join(ddir, dname + '.dem'), 'w') as f:
                f.write(text)
            # Replay: 40 shots, each with 0..6 random mechanisms (masks in 01 format).
            ne = None
            with tempfile.TemporaryDirectory() as td:
                model = os.path.join(ddir, dname + '.dem')
                # Count mechanisms via an empty replay of zero errors per shot is not possible without
                # knowing E; sample once with --err_out to learn the row width.
                run(['sample_dem', '--shots', '1', '--in', model, '--out', os.path.join(td, 'd'),
                     '--obs_out', os.path.join(td, 'o'), '--err_out', os.path.join(td, 'e'), '--seed', '1'])
                ne = len(open(os.path.join(td, 'e')).read().strip())
                masks = []
                for _ in range(40):
                    k = rng.randint(0, 6)
                    masks.append(sorted(rng.sample(range(ne), k)))
                with open(os.path.join(td, 'err.01'), 'w') as f:
                    for m in masks:
                        row = ['0'] * ne
                        for e in m:
                            row[e] = '1'
                        f.write(''.join(row) + '\n')
                run(['sample_dem', '--shots', '40', '--in', model, '--out', os.path.join(td, 'd.01'),
                     '--obs_out', os.path.join(td, 'o.01'), '--replay_err_in', os.path.join(td, 'err.01'),
                     '--replay_err_in_format', '01'])
                dets = open(os.path.join(td, 'd.01')).read().split()
                obs = open(os.path.join(td, 'o.01')).read().split()
            cases.append({'dem': dname, 'num_errors': ne, 'errors': masks, 'dets': dets, 'obs': obs})
    json.dump(cases, open(os.path.join(HERE, 'dem_replay.json'), 'w'))
    # The d=25, 25-round surface code's model with folded loops: the full-size DEM workload.
    text = run(['analyze_errors', '--in', os.path.join(HERE, 'circuits', 'surface_d25_r25.stim'),
                '--approximate_disjoint_errors', '--fold_loops'])
    with gzip.open(os.path.join(ddir, 'surface_d25_r25_folded.dem.gz'), 'wt', compresslevel=9) as f:
        f.write(text)


def ref_sample_cli(text):
    """The reference CLI's reference sample of a circuit: `stim m2d` converts an all-zero measurement
    shot of a copy of the circuit whose detectors are its single measurements; each detector is then
    the measurement's reference bit (measurements_to_detection_events.inl:75-88)."""
    body = [l for l in text.split('\n')
            if not l.strip().startswith('DETECTOR') and not l.strip().startswith('OBSERVABLE_INCLUDE')]
    m = st.stats(st.parse_flat(text))['num_measurements']
    if m == 0:
        return ''
    body += ['DETECTOR rec[-%d]' % (m - j) for j in range(m)]
    with tempfile.TemporaryDirectory() as td:
        cpath = os.path.join(td, 'c.stim')
        ipath = os.path.join(td, 'm.01')
        with open(cpath, 'w') as f:
            f.write('\n'.join(body) + '\n')
        with open(ipath, 'w') as f:
            f.write('0' * m + '\n')
        out = run(['m2d', '--circuit', cpath, '--in', ipath, '--in_format', '01', '--out_format', '01'], stdin='')
    line = out.strip()
    assert len(line) == m, (len(line), m)
    return line


def write_ref_samples():
    """Reference samples (tableau, random results taken as 0) from the reference CLI for the fixture
    circuits, gate-zoo circuits and sign-sensitive random circuits."""
    cases = []
    for name in ['surface_d3_r3', 'surface_d5_r5', 'surface_x_d5_r3', 'rep_d7_r5', 'color_d5_r3', 'unrotated_d3_r2',
                 'surface_d25_r25']:
        text = open(os.path.join(HERE, 'circuits', name + '.stim')).read()
        cases.append({'name': name, 'ref': ref_sample_cli(text)})
    for seed in range(40):
        text = zoo.gate_zoo(seed)
        cases.append({'name': 'gate_zoo_%d' % seed, 'circuit': text, 'ref': ref_sample_cli(text)})
    for seed in range(500):
        text = zoo.ref_zoo(seed, n_qubits=6 + seed % 5)
        cases.append({'name': 'ref_zoo_%d' % seed, 'circuit': text, 'ref': ref_sample_cli(text)})
    json.dump(cases, open(os.path.join(HERE, 'ref_samples.json'), 'w'))
    return cases


def write_meas_counts():
    """Per-measurement counts of ones from the reference CLI's seeded `stim sample` (reference sample
    XOR frame flips, random results randomized), for a two-sample test of the measurement sampler."""
    out = {}
    shots = 200000
    texts = {name: open(os.path.join(HERE, 'circuits', name + '.stim')).read()
             for name in ['surface_d3_r3', 'rep_d7_r5', 'color_d5_r3', 'unrotated_d3_r2', 'surface_x_d5_r3']}
    for seed in range(3):
        texts['gate_zoo_%d' % seed] = zoo.gate_zoo(seed)
        texts['ref_zoo_%d' % seed] = zoo.ref_zoo(seed)
    import numpy as np
    for name, text in texts.items():
        m = st.stats(st.parse_flat(text))['num_measurements']
        with tempfile.TemporaryDirectory() as td:
            cpath = os.path.join(td, 'c.stim')
            with open(cpath, 'w') as f:
                f.write(text)
            opath = os.path.join(td, 'o.b8')
            run(['sample', '--shots', str(shots), '--in', cpath, '--out_format', 'b8', '--out', opath, '--seed', '7'])
            raw = np.fromfile(opath, dtype=np.uint8).reshape(shots, (m + 7) // 8)
        bits = np.unpackbits(raw, axis=1, bitorder='little')[:, :m]
        out[name] = {'shots': shots, 'counts': bits.sum(axis=0).tolist(),
                     'circuit': text if name.startswith(('gate_zoo', 'ref_zoo')) else None}
    json.dump(out, open(os.path.join(HERE, 'meas_counts.json'), 'w'))


def _with_random_checks(text, rng, m):
    """Appends random detectors and observables over a circuit's measurements (the reference converter
    does not require them to be deterministic), so conversion exercises reference parities."""
    lines = [text.rstrip('\n')]
    for _ in range(rng.randint(3, 12)):
        ks = rng.sample(range(1, m + 1), k=min(m, rng.randint(1, 4)))
        lines.append('DETECTOR ' + ' '.join('rec[-%d]' % k for k in ks))
    for o in range(rng.randint(0, 3)):
        ks = rng.sample(range(1, m + 1), k=min(m, rng.randint(1, 3)))
        lines.append('OBSERVABLE_INCLUDE(%d) ' % o + ' '.join('rec[-%d]' % k for k in ks))
    return '\n'.join(lines) + '\n'


def write_m2d_fixtures():
    """Reference `stim m2d` conversions of seeded random measurement data (bit-packed rows as hex),
    with observables appended, with and without the reference sample."""
    import numpy as np
    rng = random.Random(20261204)
    nrng = np.random.default_rng(20261204)
    texts = {name: open(os.path.join(HERE, 'circuits', name + '.stim')).read()
             for name in ['surface_d3_r3', 'surface_d5_r5', 'rep_d7_r5', 'color_d5_r3', 'unrotated_d3_r2',
                          'surface_x_d5_r3', 'surface_d25_r25']}
    for seed in range(10):
        texts['gate_zoo_%d' % seed] = zoo.gate_zoo(seed)
    for seed in range(30):
        t = zoo.ref_zoo(1000 + seed, n_qubits=6 + seed % 4)
        m = st.stats(st.parse_flat(t))['num_measurements']
        texts['ref_zoo_%d' % (1000 + seed)] = _with_random_checks(t, rng, m)
    cases = []
    for name, text in texts.items():
        s = st.stats(st.parse_flat(text))
        m, nd, no = s['num_measurements'], s['num_detectors'], s['num_observables']
        shots = 16 if name == 'surface_d25_r25' else 64
        bits = nrng.random((shots, m)) < (0.5 if name.startswith(('gate', 'ref')) else 0.05)
        for skip in (False, True):
            with tempfile.TemporaryDirectory() as td:
                cpath = os.path.join(td, 'c.stim')
                with open(cpath, 'w') as f:
                    f.write(text)
                ipath = os.path.join(td, 'm.b8')
                opath = os.path.join(td, 'd.b8')
                np.packbits(bits, axis=1, bitorder='little').tofile(ipath)
                args = ['m2d', '--circuit', cpath, '--in', ipath, '--in_format', 'b8', '--out', opath,
                        '--out_format', 'b8', '--append_observables']
                if skip:
                    args.append('--skip_reference_sample')
                run(args, stdin='')
                out = np.fromfile(opath, dtype=np.uint8).reshape(shots, (nd + no + 7) // 8)
            cases.append({'name': name, 'skip_reference_sample': skip, 'shots': shots,
                          'reference_sample': ref_sample_cli(text),
                          'circuit': None if name in ('surface_d3_r3', 'surface_d5_r5', 'rep_d7_r5', 'color_d5_r3',
                                                      'unrotated_d3_r2', 'surface_x_d5_r3', 'surface_d25_r25') else text,
                          'measurements': np.packbits(bits, axis=1, bitorder='little').tobytes().hex(),
                          'expected': out.tobytes().hex()})
    json.dump(cases, open(os.path.join(HERE, 'm2d_vectors.json'), 'w'))


def write_m2d_sweep_fixtures():
    """Reference `stim m2d --sweep` conversions of seeded random measurement and sweep data on circuits
    with sweep-controlled gates (zoo.sweep_zoo), with and without the reference sample, observables
    appended (bit-packed rows as hex)."""
    import numpy as np
    nrng = np.random.default_rng(20261205)
    cases = []
    for seed in range(24):
        n_sweep = (5, 40, 70)[seed % 3]
        text = zoo.sweep_zoo(3000 + seed, n_qubits=5 + seed % 4, layers=14 + seed % 7, n_sweep=n_sweep)
        s = st.stats(st.parse_flat(text))
        m, nd, no = s['num_measurements'], s['num_detectors'], s['num_observables']
        k = max([int(x) + 1 for x in __import__('re').findall(r'sweep\[(\d+)\]', text)] + [0])
        shots = 64
        bits = nrng.random((shots, m)) < 0.5
        sweep = nrng.random((shots, k)) < 0.5
        for skip in (False, True):
            with tempfile.TemporaryDirectory() as td:
                cpath = os.path.join(td, 'c.stim')
                with open(cpath, 'w') as f:
                    f.write(text)
                ipath, spath, opath = (os.path.join(td, x) for x in ('m.b8', 's.b8', 'd.b8'))
                np.packbits(bits, axis=1, bitorder='little').tofile(ipath)
                np.packbits(sweep, axis=1, bitorder='little').tofile(spath)
                args = ['m2d', '--circuit', cpath, '--in', ipath, '--in_format', 'b8', '--sweep', spath,
                        '--sweep_format', 'b8', '--out', opath, '--out_format', 'b8', '--append_observables']
                if skip:
                    args.append('--skip_reference_sample')
                run(args, stdin='')
                out = np.fromfile(opath, dtype=np.uint8).reshape(shots, (nd + no + 7) // 8)
            cases.append({'name': 'sweep_zoo_%d' % (3000 + seed), 'skip_reference_sample': skip, 'shots': shots,
                          'circuit': text, 'num_sweep_bits': k,
                          'measurements': np.packbits(bits, axis=1, bitorder='little').tobytes().hex(),
                          'sweep': np.packbits(sweep, axis=1, bitorder='little').tobytes().hex(),
                          'expected': out.tobytes().hex()})
    json.dump(cases, open(os.path.join(HERE, 'm2d_sweep_vectors.json'), 'w'))


def write_zoo_counts():
    """Seeded detector/observable firing counts of the reference CLI (`stim detect`, 2^20 shots) on
    gate-zoo circuits whose noise includes PAULI_CHANNEL_1/2, heralded channels and correlated-error
    chains: the statistical pin of the oracle's exact rates for those channels."""
    cases = {}
    seed = 0
    while len(cases) < 6:
        text = zoo.gate_zoo(500 + seed, n_qubits=12, layers=40)
        seed += 1
        kinds = [k for k in ('PAULI_CHANNEL_1', 'PAULI_CHANNEL_2', 'HERALDED_ERASE', 'HERALDED_PAULI_CHANNEL_1',
                             'ELSE_CORRELATED_ERROR') if k in text]
        if len(kinds) < 3 or 'ELSE_CORRELATED_ERROR' not in text:
            continue
        with tempfile.NamedTemporaryFile('w', suffix='.stim', delete=False) as f:
            f.write(text)
            path = f.name
        try:
            shots = 1 << 20
            cases['zoo_counts_%d' % (500 + seed - 1)] = {'circuit': text, 'shots': shots, 'seed': 11,
                                                         'counts': seeded_counts(path, shots, 11), 'kinds': kinds}
        finally:
            os.unlink(path)
    # Large probabilities, where a chain's conditions, a heralded channel's outcome order or a
    # PAULI_CHANNEL_2 argument order change the rates by far more than the test's resolution.
    stress = '\n'.join([
        'R 0 1 2 3 4 5',
        'H 1 4',
        'E(0.25) X0 X1',
        'ELSE_CORRELATED_ERROR(0.35) X1 X2',
        'CX 4 5',
        'CX 4 5',
        'ELSE_CORRELATED_ERROR(0.45) X0 X2 Y3',
        'PAULI_CHANNEL_2(0.01, 0.02, 0.03, 0.04, 0.05, 0.06, 0.07, 0.015, 0.025, 0.035, 0.045, 0.055, 0.065, '
        '0.075, 0.085) 4 5',
        'PAULI_CHANNEL_1(0.1, 0.2, 0.05) 3',
        'HERALDED_ERASE(0.2) 2',
        'HERALDED_PAULI_CHANNEL_1(0.05, 0.1, 0.15, 0.2) 1',
        'M 0 2 3',
        'MX 1 4',
        'M 5',
        'DETECTOR rec[-8]',
        'DETECTOR rec[-7]',
    ] + ['DETECTOR rec[-%d]' % k for k in range(1, 7)]
      + ['DETECTOR rec[-%d] rec[-%d]' % (a, b) for a in range(1, 7) for b in range(a + 1, 9)]
      + ['OBSERVABLE_INCLUDE(0) rec[-6] rec[-3]']) + '\n'
    with tempfile.NamedTemporaryFile('w', suffix='.stim', delete=False) as f:
        f.write(stress)
        path = f.name
    try:
        cases['chain_stress'] = {'circuit': stress, 'shots': 1 << 20, 'seed': 11,
                                 'counts': seeded_counts(path, 1 << 20, 11),
                                 'kinds': ['PAULI_CHANNEL_1', 'PAULI_CHANNEL_2', 'HERALDED_ERASE',
                                           'HERALDED_PAULI_CHANNEL_1', 'ELSE_CORRELATED_ERROR']}
    finally:
        os.unlink(path)
    json.dump(cases, open(os.path.join(HERE, 'zoo_counts.json'), 'w'))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == '--m2d-sweep-only':
        write_m2d_sweep_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == '--zoo-counts-only':
        write_zoo_counts()
        return
    if len(sys.argv) > 1 and sys.argv[1] == '--m2d-only':
        write_m2d_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == '--meas-counts-only':
        write_meas_counts()
        return
    if len(sys.argv) > 1 and sys.argv[1] == '--ref-samples-only':
        write_ref_samples()
        return
    if len(sys.argv) > 1 and sys.argv[1] == '--gen-pins-only':
        write_gen_pins()
        return
    if len(sys.argv) > 1 and sys.argv[1] == '--dem-only':
        write_dem_fixtures()
        return
    if not os.path.isdir('/root/reference') or not os.path.exists(STIM):
        sys.exit('make_fixtures.py needs /root/reference and oracle/_ref/stim (make -C oracle ref)')
    rng = random.Random(20261202)
    cdir = os.path.join(HERE, 'circuits')
    os.makedirs(cdir, exist_ok=True)

    circuits = {
        'surface_d3_r3': gen('surface_code', 'rotated_memory_z', 3, 3, 1e-3),
        'surface_d5_r5': gen('surface_code', 'rotated_memory_z', 5, 5, 1e-3),
        'surface_d25_r25': gen('surface_code', 'rotated_memory_z', 25, 25, 1e-3),
        'surface_x_d5_r3': gen('surface_code', 'rotated_memory_x', 5, 3, 2e-3,
                               ['--before_round_data_depolarization', '0.001']),
        'unrotated_d3_r2': gen('surface_code', 'unrotated_memory_z', 3, 2, 1e-3),
        'rep_d7_r5': gen('repetition_code', 'memory', 7, 5, 1e-3),
        'color_d5_r3': gen('color_code', 'memory_xyz', 5, 3, 1e-3),
    }
    for name, text in circuits.items():
        with open(os.path.join(cdir, name + '.stim'), 'w') as f:
            f.write(text)

    write_gen_pins()
    write_dem_fixtures()

    # Fixed-fault vectors from the reference sampler.
    fixtures = []
    for name in ['surface_d3_r3', 'surface_d5_r5', 'surface_x_d5_r3', 'unrotated_d3_r2', 'rep_d7_r5', 'color_d5_r3']:
        fixtures.append(fault_fixture(name, circuits[name], 60, rng))
    for seed in range(8):
        text = zoo.gate_zoo(seed, n_qubits=6 + seed % 4, layers=8 + seed)
        fixtures.append(fault_fixture(f'zoo_{seed}', text, 40, rng))
    json.dump(fixtures, open(os.path.join(HERE, 'fault_vectors.json'), 'w'))

    # Exact per-detector rates from the reference's detector error model.
    rates = {name: dem_rates(circuits[name]) for name in ['surface_d3_r3', 'surface_d5_r5', 'surface_x_d5_r3',
                                                          'rep_d7_r5', 'color_d5_r3']}
    json.dump(rates, open(os.path.join(HERE, 'dem_rates.json'), 'w'))

    # Seeded firing counts from the reference sampler (statistical comparisons).
    counts = {}
    for name, shots in [('surface_d5_r5', 1 << 20), ('surface_d3_r3', 1 << 20)]:
        counts[name] = {'shots': shots, 'seed': 5,
                        'counts': seeded_counts(os.path.join(cdir, name + '.stim'), shots, 5)}
    json.dump(counts, open(os.path.join(HERE, 'seeded_counts.json'), 'w'))

    # Result-format files for a deterministic circuit (every shot identical).
    det_circuit = ('X_ERROR(1) 0 3 4 9\nM 0 1 2 3 4 5 6 7 8 9 10\n'
                   + ''.join(f'DETECTOR rec[-{11 - i}]\n' for i in range(11))
                   + 'OBSERVABLE_INCLUDE(0) rec[-11]\nOBSERVABLE_INCLUDE(2) rec[-8] rec[-2]\n')
    formats = {'circuit': det_circuit, 'shots': 5, 'outputs': {}}
    with tempfile.NamedTemporaryFile('w', suffix='.stim', delete=False) as f:
        f.write(det_circuit)
        path = f.name
    for fmt in ['01', 'b8', 'r8', 'hits', 'dets', 'ptb64']:
        shots = 64 if fmt == 'ptb64' else 5
        for mode in ['plain', 'append', 'prepend']:
            extra = {'plain': [], 'append': ['--append_observables'], 'prepend': ['--prepend_observables']}[mode]
            b = subprocess.run([STIM, 'detect', '--shots', str(shots), '--in', path, '--out_format', fmt] + extra,
                               capture_output=True, check=True).stdout
            formats['outputs'][f'{fmt}/{mode}/{shots}'] = b.hex()
        with tempfile.NamedTemporaryFile(suffix='.obs', delete=False) as g:
            obs_path = g.name
        r = subprocess.run([STIM, 'detect', '--shots', str(shots), '--in', path, '--out_format', fmt,
                            '--obs_out', obs_path, '--obs_out_format', fmt], capture_output=True)
        if r.returncode == 0:  # the reference aborts on some obs_out formats (e.g. dets); skip those
            formats['outputs'][f'{fmt}/obs_out_dets/{shots}'] = r.stdout.hex()
            formats['outputs'][f'{fmt}/obs_out_obs/{shots}'] = open(obs_path, 'rb').read().hex()
        os.unlink(obs_path)
    os.unlink(path)
    json.dump(formats, open(os.path.join(HERE, 'formats.json'), 'w'), indent=0)
    print('fixtures written')


if __name__ == '__main__':
    main()
