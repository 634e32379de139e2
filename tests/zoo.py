"""Seeded generator of small circuits that exercise every gate the frame path handles.

Every detector and observable it writes is deterministic in the noiseless circuit (measurement
gadgets compare repeated measurements, resets followed by measurements, feedback resets, heralds),
so a reference sampler that randomizes its frames still produces the same bits as a fixed-fault
run. Used by tests/golden/make_fixtures.py (with the reference CLI) and by the GPU parity tests.
"""

import random

ONE_Q = ['H', 'H_XY', 'H_YZ', 'H_NXY', 'H_NXZ', 'H_NYZ', 'I', 'X', 'Y', 'Z',
         'C_XYZ', 'C_NXYZ', 'C_XNYZ', 'C_XYNZ', 'C_ZYX', 'C_ZYNX', 'C_ZNYX', 'C_NZYX',
         'SQRT_X', 'SQRT_X_DAG', 'SQRT_Y', 'SQRT_Y_DAG', 'S', 'S_DAG', 'SQRT_Z', 'SQRT_Z_DAG', 'H_XZ']
TWO_Q = ['XCX', 'XCY', 'XCZ', 'YCX', 'YCY', 'YCZ', 'CX', 'CY', 'CZ', 'CNOT', 'ZCX', 'ZCY', 'ZCZ', 'II',
         'SQRT_XX', 'SQRT_XX_DAG', 'SQRT_YY', 'SQRT_YY_DAG', 'SQRT_ZZ', 'SQRT_ZZ_DAG',
         'SWAP', 'ISWAP', 'ISWAP_DAG', 'CXSWAP', 'SWAPCX', 'CZSWAP', 'SWAPCZ']


def _p(rng):
    return round(rng.uniform(0.001, 0.02), 4)


def _noise(rng, qs):
    kind = rng.randrange(9)
    a = rng.sample(qs, k=min(len(qs), rng.randint(1, 4)))
    if kind == 0:
        return f'X_ERROR({_p(rng)}) ' + ' '.join(map(str, a))
    if kind == 1:
        return f'Y_ERROR({_p(rng)}) ' + ' '.join(map(str, a))
    if kind == 2:
        return f'Z_ERROR({_p(rng)}) ' + ' '.join(map(str, a))
    if kind == 3:
        return f'DEPOLARIZE1({_p(rng)}) ' + ' '.join(map(str, a))
    if kind == 4:
        return f'PAULI_CHANNEL_1({_p(rng)}, {_p(rng)}, {_p(rng)}) ' + ' '.join(map(str, a))
    if kind in (5, 6):
        b = rng.sample(qs, k=2 * (len(qs) // 2))
        b = b[:2 * rng.randint(1, max(1, len(b) // 2))]
        if kind == 5:
            return f'DEPOLARIZE2({_p(rng)}) ' + ' '.join(map(str, b))
        args = ', '.join(str(round(rng.uniform(0.0005, 0.003), 5)) for _ in range(15))
        return f'PAULI_CHANNEL_2({args}) ' + ' '.join(map(str, b))
    # correlated chain
    lines = []
    for link in range(rng.randint(1, 3)):
        name = 'E' if link == 0 else 'ELSE_CORRELATED_ERROR'
        ps = rng.sample(qs, k=rng.randint(1, 3))
        lines.append(f'{name}({_p(rng)}) ' + ' '.join(rng.choice('XYZ') + str(q) for q in ps))
    return '\n'.join(lines)


def _product(rng, qs, w):
    ps = rng.sample(qs, k=w)
    return '*'.join(rng.choice('XYZ') + str(q) for q in ps)


def _gadget(rng, qs, nobs):
    """Returns (lines, observable_lines) producing deterministic detectors."""
    k = rng.randrange(12)
    q = rng.choice(qs)
    if k == 0:
        w = rng.randint(1, min(4, len(qs)))
        P = _product(rng, qs, w)
        inv = '!' if rng.random() < 0.3 else ''
        lines = [f'MPP({_p(rng)}) {inv}{P}', f'MPP {P}', 'DETECTOR rec[-1] rec[-2]']
        if nobs and rng.random() < 0.5:
            lines.append(f'OBSERVABLE_INCLUDE({rng.randrange(nobs)}) rec[-1] rec[-2]')
        return lines
    if k == 1:
        m = rng.choice(['M', 'MX', 'MY', 'MZ'])
        return [f'{m}({_p(rng)}) {q}', f'{m} {q}', 'DETECTOR rec[-1] rec[-2]']
    if k == 2:
        mr, m = rng.choice([('MR', 'M'), ('MRX', 'MX'), ('MRY', 'MY'), ('MRZ', 'MZ')])
        return [f'{mr}({_p(rng)}) {q}', f'{m} {q}', 'DETECTOR rec[-1]']
    if k == 3:
        r, m = rng.choice([('R', 'M'), ('RX', 'MX'), ('RY', 'MY'), ('RZ', 'MZ')])
        return [f'{r} {q}', f'{m} {q}', 'DETECTOR rec[-1]']
    if k == 4:
        g = rng.choice(['CX', 'CNOT', 'ZCX'])
        return [f'M {q}', f'{g} rec[-1] {q}', f'M {q}', 'DETECTOR rec[-1]']
    if k == 5:
        # MX then CZ rec feedback resets the qubit to |+>.
        g = rng.choice(['CZ', 'ZCZ'])
        order = f'{g} rec[-1] {q}' if rng.random() < 0.5 else f'{g} {q} rec[-1]'
        return [f'MX {q}', order, f'MX {q}', 'DETECTOR rec[-1]']
    if k == 6:
        # MY then CY rec feedback... Y-basis feedback uses a Y flip after a Z measurement.
        return [f'M {q}', f'CY rec[-1] {q}', f'M {q}', 'DETECTOR rec[-1]']
    if k == 7 and len(qs) >= 2:
        a, b = rng.sample(qs, 2)
        g = rng.choice(['MXX', 'MYY', 'MZZ'])
        lines = [f'{g}({_p(rng)}) {a} {b}', f'{g} {a} {b}', 'DETECTOR rec[-1] rec[-2]']
        if nobs and rng.random() < 0.5:
            lines.append(f'OBSERVABLE_INCLUDE({rng.randrange(nobs)}) rec[-2] rec[-1]')
        return lines
    if k == 8:
        return [f'HERALDED_ERASE({_p(rng)}) {q}', 'DETECTOR rec[-1]']
    if k == 9:
        return [f'HERALDED_PAULI_CHANNEL_1({_p(rng)}, {_p(rng)}, {_p(rng)}, {_p(rng)}) {q}', 'DETECTOR rec[-1]']
    if k == 10:
        return [f'MPAD({_p(rng)}) 0 1', 'DETECTOR rec[-1]', 'DETECTOR rec[-2]']
    # sweep-controlled gates are no-ops without a sweep table
    return [f'CX sweep[{rng.randrange(4)}] {q}', f'CZ {q} sweep[0]']


def gate_zoo(seed, n_qubits=8, layers=10, nobs=2, repeat=True):
    """A random circuit touching every frame rule, noise channel and measurement gadget."""
    rng = random.Random(seed)
    qs = list(range(n_qubits))
    out = ['R ' + ' '.join(map(str, qs))]

    def layer():
        lines = []
        k1 = rng.sample(qs, k=rng.randint(1, n_qubits))
        for q in k1:
            lines.append(f'{rng.choice(ONE_Q)} {q}')
        perm = rng.sample(qs, k=n_qubits)
        pairs = [(perm[2 * i], perm[2 * i + 1]) for i in range(n_qubits // 2)]
        g = rng.choice(TWO_Q)
        lines.append(g + ' ' + ' '.join(f'{a} {b}' for a, b in pairs[:rng.randint(1, len(pairs))]))
        if rng.random() < 0.3:
            w = rng.randint(1, 3)
            lines.append(('SPP ' if rng.random() < 0.5 else 'SPP_DAG ') + _product(rng, qs, w))
        lines.append(_noise(rng, qs))
        lines.extend(_gadget(rng, qs, nobs))
        lines.append('TICK')
        return lines

    for i in range(layers):
        out.extend(layer())
        if repeat and i == layers // 2:
            out.append(f'REPEAT {rng.randint(2, 3)} {{')
            for _ in range(2):
                out.extend('    ' + line for line in '\n'.join(layer()).split('\n'))
            out.append('}')
    # Frame-dependent observable: measure everything in Z after undoing nothing is not deterministic,
    # so Pauli-target observables are only added on freshly reset qubits.
    q = rng.choice(qs)
    out.append(f'RX {q}')
    out.append(f'OBSERVABLE_INCLUDE({nobs - 1 if nobs else 0}) X{q}')
    return '\n'.join(out) + '\n'


def ref_zoo(seed, n_qubits=6, layers=14):
    """Seeded circuits whose reference sample depends on the sign conventions of every gate:
    random Clifford layers, measurements and resets in every basis (some results inverted),
    Pauli-product measurements and phases (repeated factors, inverted factors, measurement bits in
    SPP products), pair measurements, feedback from measurement records, MPAD, and noise (which the
    reference sample ignores; heralded errors record 0). Ends by measuring every qubit."""
    rng = random.Random(seed)
    qs = list(range(n_qubits))
    lines = ['RX 0', 'RY 1']
    nmeas = 0

    def sub(k=None):
        return rng.sample(qs, k=k or rng.randint(1, len(qs)))

    def inv(q):
        return ('!' if rng.random() < 0.3 else '') + str(q)

    def product():
        ps = rng.sample(qs, k=rng.randint(1, 3))
        fs = [rng.choice('XYZ') + str(q) for q in ps]
        if rng.random() < 0.3:
            fs.append(fs[0])            # a repeated factor cancels
        if rng.random() < 0.2:
            q = ps[0]
            fs += ['X%d' % q, 'Z%d' % q, 'X%d' % q, 'Z%d' % q]   # XZXZ = -I
        return '*'.join(('!' if rng.random() < 0.25 else '') + f for f in fs)

    for _ in range(layers):
        k = rng.randrange(12)
        if k < 3:
            lines.append(rng.choice(ONE_Q) + ' ' + ' '.join(map(str, sub())))
        elif k < 5:
            ps = sub(2 * rng.randint(1, len(qs) // 2))
            lines.append(rng.choice(TWO_Q) + ' ' + ' '.join(map(str, ps)))
        elif k == 5:
            ts = sub()
            name = rng.choice(['M', 'MX', 'MY', 'MR', 'MRX', 'MRY'])
            arg = '(0.05)' if rng.random() < 0.3 else ''
            lines.append(name + arg + ' ' + ' '.join(inv(q) for q in ts))
            nmeas += len(ts)
        elif k == 6:
            lines.append(rng.choice(['R', 'RX', 'RY']) + ' ' + ' '.join(map(str, sub())))
        elif k == 7:
            n = rng.randint(1, 3)
            lines.append('MPP ' + ' '.join(product() for _ in range(n)))
            nmeas += n
        elif k == 8:
            lines.append(rng.choice(['SPP', 'SPP_DAG']) + ' ' + ' '.join(product() for _ in range(rng.randint(1, 2))))
        elif k == 9 and nmeas:
            q = rng.choice(qs)
            r = 'rec[-%d]' % rng.randint(1, min(nmeas, 6))
            lines.append(rng.choice(['CX %s %d', 'CY %s %d', 'CZ %s %d']) % (r, q) if rng.random() < 0.6
                         else rng.choice(['XCZ %d %s', 'YCZ %d %s', 'CZ %d %s']) % (q, r))
        elif k == 10:
            pairs = []
            for _ in range(rng.randint(1, 3)):
                a, b = rng.sample(qs, 2)
                pairs.append('%s %s' % (inv(a), inv(b)))
            lines.append('%s %s' % (rng.choice(['MXX', 'MYY', 'MZZ']), ' '.join(pairs)))
            nmeas += len(pairs)
        else:
            kind = rng.randrange(4)
            if kind == 0:
                lines.append('MPAD %d %d' % (rng.randint(0, 1), rng.randint(0, 1)))
                nmeas += 2
            elif kind == 1:
                lines.append('HERALDED_ERASE(0.2) %d' % rng.choice(qs))
                nmeas += 1
            elif kind == 2:
                lines.append('DEPOLARIZE1(0.3) ' + ' '.join(map(str, sub())))
            else:
                lines.append('X_ERROR(0.4) ' + ' '.join(map(str, sub())))
    lines.append(rng.choice(['M', 'MX', 'MY']) + ' ' + ' '.join(inv(q) for q in qs))
    return '\n'.join(lines) + '\n'


def sweep_zoo(seed, n_qubits=6, layers=16, n_sweep=5):
    """Seeded circuits whose detection events depend on sweep data: sweep-controlled CX/CY/CZ/XCZ/YCZ
    (sweep bit as control, or on either side of CZ; mixed with quantum and record-controlled pairs in
    one instruction), random Clifford layers, measurements and resets in every basis, MPP, feedback, a
    REPEAT block, noise (ignored by the conversion) and detectors / observables over random records,
    including observables with Pauli targets (which sweep-driven frame flips reach)."""
    rng = random.Random(seed)
    qs = list(range(n_qubits))
    lines = ['R ' + ' '.join(map(str, qs))]
    nmeas = 0

    def sweep_pair():
        q, k = rng.choice(qs), rng.randrange(n_sweep)
        g = rng.choice(['CX', 'CY', 'CZ', 'CZ', 'XCZ', 'YCZ'])
        if g in ('XCZ', 'YCZ'):
            return g, f'{q} sweep[{k}]'
        if g == 'CZ' and rng.random() < 0.5:
            return g, f'{q} sweep[{k}]'
        return g, f'sweep[{k}] {q}'

    def layer():
        nonlocal nmeas
        out = []
        k = rng.randrange(9)
        if k < 2:
            out.append(rng.choice(ONE_Q) + ' ' + ' '.join(map(str, rng.sample(qs, rng.randint(1, len(qs))))))
        elif k < 4:
            ps = rng.sample(qs, 2 * rng.randint(1, len(qs) // 2))
            out.append(rng.choice(TWO_Q) + ' ' + ' '.join(map(str, ps)))
        elif k < 6:
            g, pair = sweep_pair()
            extra = ''
            if rng.random() < 0.4:
                a, b = rng.sample(qs, 2)
                extra = f' {a} {b}' if rng.random() < 0.5 else f' sweep[{rng.randrange(n_sweep)}] {a}' \
                    if g in ('CX', 'CY', 'CZ') else ''
            out.append(f'{g} {pair}{extra}')
        elif k == 6:
            ts = rng.sample(qs, rng.randint(1, len(qs)))
            out.append(rng.choice(['M', 'MX', 'MY', 'MR', 'MRX', 'MRY']) + ' ' + ' '.join(map(str, ts)))
            nmeas += len(ts)
        elif k == 7:
            out.append('MPP ' + _product(rng, qs, rng.randint(1, 3)))
            nmeas += 1
        else:
            out.append(_noise(rng, qs))
        if nmeas and rng.random() < 0.5:
            ks = rng.sample(range(1, nmeas + 1), k=min(nmeas, rng.randint(1, 3)))
            out.append('DETECTOR ' + ' '.join('rec[-%d]' % k for k in ks))
        if nmeas and rng.random() < 0.15:
            out.append(f'CX rec[-{rng.randint(1, min(nmeas, 3))}] {rng.choice(qs)}')
        return out

    for i in range(layers):
        lines.extend(layer())
        if i == layers // 2:
            lines.append('REPEAT 2 {')
            for _ in range(3):
                lines.extend('    ' + x for x in layer())
            lines.append('}')
            nmeas += 0
    ts = list(qs)
    lines.append('M ' + ' '.join(map(str, ts)))
    nmeas += len(ts)
    for q in qs[:3]:
        lines.append(f'DETECTOR rec[-{n_qubits - q}]')
    lines.append('OBSERVABLE_INCLUDE(0) rec[-1] ' + rng.choice('XYZ') + str(rng.choice(qs)))
    lines.append('OBSERVABLE_INCLUDE(1) ' + rng.choice('XYZ') + str(rng.choice(qs)))
    return '\n'.join(lines) + '\n'
