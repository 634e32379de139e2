"""Workload generators for the sampler's benchmark and serving configurations.

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


def _fmt(p):
    return repr(float(p))


def random_clifford(n_qubits=10000, layers=1000, seed=0, px=3e-5, py=3e-5, pz=3e-5, p2=1e-4, block=10,
                    n_products=500, measured_fraction=0.1):
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
            out.append('MPP ' + ' '.join(prods))
            out.append('MPP ' + ' '.join(prods))
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
