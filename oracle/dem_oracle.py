"""Test infrastructure only: a CPU restatement of the reference's detector-error-model sampler.

Used by tests/ as the checker for the GPU DEM sampler; never by the product path.

Follows:
  - doc/file_format_dem_detector_error_model.md (syntax: instructions, tags, args, targets, repeat
    blocks; semantics: shift_detectors offsets relative detector targets, `^` only separates a
    suggested decomposition, repeated targets cancel);
  - src/stim/simulators/dem_sampler.inl:50-77 (each error instruction, REPEAT-unrolled, is one
    mechanism; it fires independently with its probability and XORs its targets into the outputs).

Pinned by: the reference CLI's `sample_dem --replay_err_in` outputs (tests/golden/dem_replay.json,
bit-exact, including a model with a folded REPEAT block); exact per-detector rates of unfolded and
folded models of the same circuit must agree, and must equal the circuit's exact rates
(tests/golden/dem_rates.json).
"""

import math

import numpy as np


def parse(text):
    """Returns (num_detectors, num_observables, [(p, dets, obs), ...]) with REPEAT blocks unrolled."""
    lines = []
    for raw in text.split('\n'):
        t = raw.split('#', 1)[0].strip()
        if t:
            lines.append(t)
    pos = 0

    def block():
        nonlocal pos
        body = []
        while pos < len(lines):
            t = lines[pos]
            pos += 1
            if t == '}':
                return body
            opens = t.endswith('{')
            if opens:
                t = t[:-1].strip()
            name_end = 0
            while name_end < len(t) and (t[name_end].isalnum() or t[name_end] == '_'):
                name_end += 1
            name = t[:name_end].lower()
            rest = t[name_end:]
            if rest.startswith('['):
                rest = rest[rest.index(']') + 1:]
            args = []
            if rest.startswith('('):
                close = rest.index(')')
                args = [float(a) for a in rest[1:close].split(',')]
                rest = rest[close + 1:]
            targets = rest.split()
            if name == 'repeat':
                body.append(('repeat', int(targets[0]), block()))
            else:
                body.append((name, args, targets))
        return body

    prog = block()
    nd = 0
    no = 0
    errors = []
    offset = 0

    def run(b):
        nonlocal nd, no, offset
        for ins in b:
            if ins[0] == 'repeat':
                for _ in range(ins[1]):
                    run(ins[2])
                continue
            name, args, targets = ins
            if name == 'error':
                dets, obs = {}, {}
                for t in targets:
                    if t[0] in 'Dd':
                        d = offset + int(t[1:])
                        dets[d] = dets.get(d, 0) ^ 1
                        nd = max(nd, d + 1)
                    elif t[0] in 'Ll':
                        k = int(t[1:])
                        obs[k] = obs.get(k, 0) ^ 1
                        no = max(no, k + 1)
                errors.append((args[0], sorted(d for d, v in dets.items() if v), sorted(k for k, v in obs.items() if v)))
            elif name == 'detector':
                for t in targets:
                    nd = max(nd, offset + int(t[1:]) + 1)
            elif name == 'logical_observable':
                for t in targets:
                    no = max(no, int(t[1:]) + 1)
            elif name == 'shift_detectors':
                offset += sum(int(t) for t in targets)

    run(prog)
    return nd, no, errors


def replay(model, masks):
    """Detection events and observable flips for per-shot lists of fired mechanism indices."""
    nd, no, errors = model
    dets = np.zeros((len(masks), nd), dtype=bool)
    obs = np.zeros((len(masks), no), dtype=bool)
    for s, fired in enumerate(masks):
        for e in fired:
            _, ed, eo = errors[e]
            dets[s, ed] ^= True
            obs[s, eo] ^= True
    return dets, obs


def exact_rates(model):
    """Firing probability of every detector then observable: P(odd number of its mechanisms fire)
    = (1 - prod(1 - 2 p_e)) / 2 over independent mechanisms."""
    nd, no, errors = model
    logs = np.zeros(nd + no)
    zero = np.zeros(nd + no, dtype=bool)   # a mechanism with p = 1/2 makes its targets uniform
    for p, ed, eo in errors:
        f = 1 - 2 * p
        for i in list(ed) + [nd + k for k in eo]:
            if f == 0:
                zero[i] = True
            else:
                logs[i] += math.log(abs(f))
    prod = np.exp(logs)
    signs = np.ones(nd + no)
    for p, ed, eo in errors:
        if p > 0.5:
            for i in list(ed) + [nd + k for k in eo]:
                signs[i] = -signs[i]
    return np.where(zero, 0.5, (1 - signs * prod) / 2)
