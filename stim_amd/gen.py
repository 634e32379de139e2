"""Workload generators for the sampler's benchmark and serving configurations.

generate(): the reference's circuit generators for the families the benchmarks use -- rotated and
unrotated surface-code memory experiments and the repetition-code memory experiment -- producing the
same circuit text as the reference's `stim gen` / `stim.Circuit.generated` (gen/gen_surface_code.cc,
gen/gen_rep_code.cc, gen/circuit_gen_params.cc), without the comment header.

random_clifford(): seeded random-Clifford stress circuits. Each layer applies a random single-qubit
Clifford from {H, S, S_DAG, SQRT_X} to every qubit and one of {CX, CZ, SWAP, ISWAP} to each pair of a
random perfect matching, with PAULI_CHANNEL_1 on every qubit and DEPOLARIZE2 on every pair. Every
`block` layers it measures `n_products` random weight-4 Pauli products on disjoint qubit sets twice in
a row (MPP), compares each pair with a DETECTOR (deterministic: the products commute and nothing
acts between the two measurements), measures 10% of the qubits with M / MR, and resets the
M-measured qubits by classical feedback (CX rec[-k] q).

surface_family(): rotated surface-code memory circuits for a set of distances and noise strengths,
obtained by substituting the probability into a template circuit of the same distance and rounds.
"""

import random


class _Builder:
    """Instruction list with the reference's fusion rule: an instruction joins the previous one when
    both have the same gate and arguments and the gate is fusable (circuit_instruction.cc:307-310;
    annotations, REPEAT and correlated errors are not)."""

    NOT_FUSABLE = {'DETECTOR', 'OBSERVABLE_INCLUDE', 'QUBIT_COORDS', 'SHIFT_COORDS', 'MPAD', 'REPEAT', 'E',
                   'ELSE_CORRELATED_ERROR'}

    def __init__(self):
        self.ops = []   # [name, args tuple, targets list] or ['REPEAT', count, _Builder]

    def add(self, name, targets=(), args=()):
        targets = list(targets)
        if name not in ('TICK',) and not targets and name not in ('SHIFT_COORDS',):
            return self
        if (self.ops and self.ops[-1][0] == name and tuple(self.ops[-1][1]) == tuple(args)
                and name not in self.NOT_FUSABLE):
            self.ops[-1][2].extend(targets)
        else:
            self.ops.append([name, tuple(args), targets])
        return self

    def extend(self, other):
        # Concatenation fuses the boundary pair only (circuit.cc:592-617).
        for k, op in enumerate(other.ops):
            if k == 0 and op[0] != 'REPEAT':
                self.add(op[0], op[2], op[1])
            else:
                self.ops.append([op[0], op[1], list(op[2]) if op[0] != 'REPEAT' else op[2]])
        return self

    def repeated(self, n):
        out = _Builder()
        if n == 1:
            out.extend(self)
        elif n > 1:
            out.ops.append(['REPEAT', n, self])
        return out

    def lines(self, indent=''):
        for name, args, targets in self.ops:
            if name == 'REPEAT':
                yield f'{indent}REPEAT {args} {{'
                yield from targets.lines(indent + '    ')
                yield indent + '}'
                continue
            head = name
            if args:
                head += '(' + ', '.join(_num(a) for a in args) + ')'
            yield indent + head + ''.join(' ' + _target(t) for t in targets)

    def text(self):
        return '\n'.join(self.lines()) + '\n'


def _num(x):
    """Formats an argument like a C++ ostream at default precision (six significant digits),
    integers without a decimal point (circuit_instruction.cc:339-355)."""
    x = float(x)
    if x == int(x) and abs(x) < 2 ** 63:
        return str(int(x))
    return format(x, 'g')


def _target(t):
    return f'rec[-{t[1]}]' if isinstance(t, tuple) else str(t)


def _rec(k):
    return ('rec', k)


class _Params:
    """Noise parameters of a generated circuit and the reference's helpers that place them
    (circuit_gen_params.cc:7-96)."""

    def __init__(self, after_clifford_depolarization=0.0, before_round_data_depolarization=0.0,
                 before_measure_flip_probability=0.0, after_reset_flip_probability=0.0):
        self.acd = float(after_clifford_depolarization)
        self.brd = float(before_round_data_depolarization)
        self.bmf = float(before_measure_flip_probability)
        self.arf = float(after_reset_flip_probability)
        for name, v in [('before_measure_flip_probability', self.bmf),
                        ('before_round_data_depolarization', self.brd),
                        ('after_clifford_depolarization', self.acd),
                        ('after_reset_flip_probability', self.arf)]:
            if not 0 <= v <= 1:
                raise ValueError(f'not 0 <= {name} <= 1')

    @staticmethod
    def _anti_basis(c, targets, p, basis):
        if p > 0:
            c.add('Z_ERROR' if basis == 'X' else 'X_ERROR', targets, (p,))

    def begin_round(self, c, data):
        c.add('TICK')
        if self.brd > 0:
            c.add('DEPOLARIZE1', data, (self.brd,))

    def unitary1(self, c, name, targets):
        c.add(name, targets)
        if self.acd > 0:
            c.add('DEPOLARIZE1', targets, (self.acd,))

    def unitary2(self, c, name, targets):
        c.add(name, targets)
        if self.acd > 0:
            c.add('DEPOLARIZE2', targets, (self.acd,))

    def reset(self, c, targets, basis='Z'):
        c.add('R' + basis if basis != 'Z' else 'R', targets)
        self._anti_basis(c, targets, self.arf, basis)

    def measure(self, c, targets, basis='Z'):
        self._anti_basis(c, targets, self.bmf, basis)
        c.add('M' + basis if basis != 'Z' else 'M', targets)

    def measure_reset(self, c, targets, basis='Z'):
        self._anti_basis(c, targets, self.bmf, basis)
        c.add('MR' + basis if basis != 'Z' else 'MR', targets)
        self._anti_basis(c, targets, self.arf, basis)


def _surface(d, rounds, params, rotated, memory_x):
    """Surface-code memory experiment (gen_surface_code.cc:32-335). Rotated-layout coordinates are
    the reference's: data qubits at odd (x, y), measurement qubits at even (x, y)."""
    data, xm, zm = set(), set(), set()
    x_obs, z_obs = [], []
    if rotated:
        for i in range(d):
            for j in range(d):
                q = (2 * i + 1, 2 * j + 1)
                data.add(q)
                if j == 0:
                    z_obs.append(q)
                if i == 0:
                    x_obs.append(q)
        for x in range(d + 1):
            for y in range(d + 1):
                parity = (x % 2) != (y % 2)
                if (x in (0, d)) and parity:
                    continue
                if (y in (0, d)) and not parity:
                    continue
                (xm if parity else zm).add((2 * x, 2 * y))
        z_order = [(1, 1), (1, -1), (-1, 1), (-1, -1)]
        x_order = [(1, 1), (-1, 1), (1, -1), (-1, -1)]

        def index(q):
            # The reference's coordinate-to-index map: x + (y - x mod 2) * (d + 1/2).
            x, y = q
            return x + (y - x % 2) * (2 * d + 1) // 2
    else:
        n = 2 * d - 1
        for x in range(n):
            for y in range(n):
                if (x % 2) != (y % 2):
                    (zm if x % 2 == 0 else xm).add((x, y))
                else:
                    data.add((x, y))
                    if x == 0:
                        x_obs.append((x, y))
                    if y == 0:
                        z_obs.append((x, y))
        z_order = x_order = [(1, 0), (0, 1), (0, -1), (-1, 0)]

        def index(q):
            return q[0] + q[1] * n
    obs = x_obs if memory_x else z_obs
    basis_meas = xm if memory_x else zm
    p2q = {}
    for q in list(data) + list(xm) + list(zm):
        p2q[q] = index(q)
    q2p = {v: k for k, v in p2q.items()}
    data_q = sorted(p2q[q] for q in data)
    x_q = sorted(p2q[q] for q in xm)
    meas_q = sorted([p2q[q] for q in xm] + [p2q[q] for q in zm])
    data_order = {q2p[q]: i for i, q in enumerate(data_q)}
    meas_order = {q2p[q]: i for i, q in enumerate(meas_q)}
    cx = [[], [], [], []]
    for k in range(4):
        for m in sorted(xm):
            dq = (m[0] + x_order[k][0], m[1] + x_order[k][1])
            if dq in p2q:
                cx[k] += [p2q[m], p2q[dq]]
        for m in sorted(zm):
            dq = (m[0] + z_order[k][0], m[1] + z_order[k][1])
            if dq in p2q:
                cx[k] += [p2q[dq], p2q[m]]
    cycle = _Builder()
    params.begin_round(cycle, data_q)
    params.unitary1(cycle, 'H', x_q)
    for t in cx:
        cycle.add('TICK')
        params.unitary2(cycle, 'CX', t)
    cycle.add('TICK')
    params.unitary1(cycle, 'H', x_q)
    cycle.add('TICK')
    params.measure_reset(cycle, meas_q)

    def coords(q):
        return q
    head = _Builder()
    for qi in sorted(q2p):
        head.add('QUBIT_COORDS', [qi], coords(q2p[qi]))
    params.reset(head, data_q, 'X' if memory_x else 'Z')
    params.reset(head, meas_q)
    head.extend(cycle)
    nm = len(meas_q)
    for m in sorted(basis_meas):
        head.add('DETECTOR', [_rec(nm - meas_order[m])], coords(m) + (0,))
    body = _Builder().extend(cycle)
    body.add('SHIFT_COORDS', [], (0, 0, 1))
    for mq in meas_q:
        m = q2p[mq]
        k = nm - meas_order[m] - 1
        body.add('DETECTOR', [_rec(k + 1), _rec(k + 1 + nm)], coords(m) + (0,))
    tail = _Builder()
    params.measure(tail, data_q, 'X' if memory_x else 'Z')
    nd = len(data_q)
    for m in sorted(basis_meas):
        ks = []
        for delta in z_order:
            dq = (m[0] + delta[0], m[1] + delta[1])
            if dq in p2q:
                ks.append(nd - data_order[dq])
        ks.append(nd + nm - meas_order[m])
        tail.add('DETECTOR', [_rec(k) for k in sorted(ks)], coords(m) + (1,))
    tail.add('OBSERVABLE_INCLUDE', [_rec(k) for k in sorted(nd - data_order[q] for q in obs)], (0,))
    out = _Builder().extend(head).extend(body.repeated(rounds - 1)).extend(tail)
    return out.text()


def _repetition(d, rounds, params):
    """Repetition-code memory experiment (gen_rep_code.cc:15-97)."""
    m = d - 1
    n = 2 * m + 1
    data_q = list(range(0, n, 2))
    meas_q = list(range(1, n, 2))
    t1, t2 = [], []
    for q in meas_q:
        t1 += [q - 1, q]
        t2 += [q + 1, q]
    cycle = _Builder()
    params.begin_round(cycle, data_q)
    params.unitary2(cycle, 'CX', t1)
    cycle.add('TICK')
    params.unitary2(cycle, 'CX', t2)
    cycle.add('TICK')
    params.measure_reset(cycle, meas_q)
    head = _Builder()
    params.reset(head, list(range(n)))
    head.extend(cycle)
    for k in range(m):
        head.add('DETECTOR', [_rec(m - k)], (2 * k + 1, 0))
    body = _Builder().extend(cycle)
    body.add('SHIFT_COORDS', [], (0, 1))
    for k in range(m):
        body.add('DETECTOR', [_rec(m - k), _rec(2 * m - k)], (2 * k + 1, 0))
    tail = _Builder()
    params.measure(tail, data_q)
    for k in range(m):
        tail.add('DETECTOR', [_rec(m - k), _rec(m - k + 1), _rec(2 * m - k + 1)], (2 * k + 1, 1))
    tail.add('OBSERVABLE_INCLUDE', [_rec(1)], (0,))
    return _Builder().extend(head).extend(body.repeated(rounds - 1)).extend(tail).text()


def generate(code_task, *, distance, rounds, after_clifford_depolarization=0.0,
             before_round_data_depolarization=0.0, before_measure_flip_probability=0.0,
             after_reset_flip_probability=0.0):
    """Circuit text of `stim.Circuit.generated(code_task, ...)` for 'surface_code:rotated_memory_z',
    'surface_code:rotated_memory_x', 'surface_code:unrotated_memory_z',
    'surface_code:unrotated_memory_x' and 'repetition_code:memory'
    (circuit.pybind.cc:1337-1383; gen/circuit_gen_params.cc:17-30 for the checks)."""
    params = _Params(after_clifford_depolarization, before_round_data_depolarization,
                     before_measure_flip_probability, after_reset_flip_probability)
    if rounds < 1:
        raise ValueError('Need rounds >= 1.')
    if distance < 2:
        raise ValueError('Need a distance >= 2.')
    code, _, task = code_task.partition(':')
    if code == 'surface_code':
        table = {'rotated_memory_z': (True, False), 'rotated_memory_x': (True, True),
                 'unrotated_memory_z': (False, False), 'unrotated_memory_x': (False, True)}
        if task not in table:
            raise ValueError(f"Unrecognized task '{task}'.")
        return _surface(distance, rounds, params, *table[task])
    if code == 'repetition_code':
        if task != 'memory':
            raise ValueError(f"Unrecognized task '{task}'.")
        return _repetition(distance, rounds, params)
    raise ValueError(f"Unrecognized circuit type '{code}'.")


def _fmt(p):
    return repr(float(p))


def random_clifford(n_qubits=10000, layers=1000, seed=0, px=3e-5, py=3e-5, pz=3e-5, p2=1e-4, block=10,
                    n_products=500, measured_fraction=0.1, mpp_flip=0.0):
    """mpp_flip > 0 makes the product measurements noisy (MPP(p)); without it the comparing detectors
    are noiseless, since nothing acts between the two measurements of a product."""
    if n_qubits < 8 or n_qubits % 2:
        raise ValueError('n_qubits must be an even number >= 8')
    rng = random.Random(seed)
    qs = list(range(n_qubits))
    out = []
    one_q = ['H', 'S', 'S_DAG', 'SQRT_X']
    two_q = ['CX', 'CZ', 'SWAP', 'ISWAP']
    n_products = min(n_products, n_qubits // 4)
    observable_done = False
    for layer in range(layers):
        groups = {g: [] for g in one_q}
        for q in qs:
            groups[rng.choice(one_q)].append(q)
        for g in one_q:
            if groups[g]:
                out.append(g + ' ' + ' '.join(map(str, groups[g])))
        out.append(f'PAULI_CHANNEL_1({_fmt(px)}, {_fmt(py)}, {_fmt(pz)}) ' + ' '.join(map(str, qs)))
        perm = qs[:]
        rng.shuffle(perm)
        pairs = {g: [] for g in two_q}
        for i in range(0, n_qubits, 2):
            pairs[rng.choice(two_q)].extend((perm[i], perm[i + 1]))
        for g in two_q:
            if pairs[g]:
                out.append(g + ' ' + ' '.join(map(str, pairs[g])))
        out.append(f'DEPOLARIZE2({_fmt(p2)}) ' + ' '.join(map(str, perm)))
        out.append('TICK')
        if (layer + 1) % block == 0:
            chosen = rng.sample(qs, 4 * n_products)
            prods = []
            for k in range(n_products):
                prods.append('*'.join(rng.choice('XYZ') + str(q) for q in chosen[4 * k:4 * k + 4]))
            mpp = f'MPP({_fmt(mpp_flip)}) ' if mpp_flip > 0 else 'MPP '
            out.append(mpp + ' '.join(prods))
            out.append(mpp + ' '.join(prods))
            for k in range(n_products):
                out.append(f'DETECTOR rec[-{n_products - k}] rec[-{2 * n_products - k}]')
            if not observable_done:
                out.append(f'OBSERVABLE_INCLUDE(0) rec[-{n_products}] rec[-{2 * n_products}]')
                observable_done = True
            nm = max(2, int(n_qubits * measured_fraction) // 2 * 2)
            meas = rng.sample(qs, nm)
            m_part, mr_part = meas[:nm // 2], meas[nm // 2:]
            out.append('M ' + ' '.join(map(str, m_part)))
            out.append('MR ' + ' '.join(map(str, mr_part)))
            # Feedback: flip each M-measured qubit back to |0> when its result was 1.
            fb = []
            for i, q in enumerate(m_part):
                fb.append(f'rec[-{nm - i}] {q}')
            out.append('CX ' + ' '.join(fb))
            out.append('TICK')
    return '\n'.join(out) + '\n'


def substitute_probability(template_text, old, new):
    """Replaces the noise strength `old` (as printed in a generated circuit) by `new`."""
    return template_text.replace(f'({old})', f'({_fmt(new)})')


def log_uniform(rng, lo, hi):
    import math
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))
