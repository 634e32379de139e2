"""CPU restatement (numpy) of the reference's Pauli-frame detection-event sampler.

TEST INFRASTRUCTURE ONLY. This module is the checker: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import it. The product (stim_amd) never routes through it.

It restates, one frame column per shot (numpy bool rows, shots along the minor axis):
  - FrameSimulator::do_* handlers           src/stim/simulators/frame_simulator.inl:139-1113
  - reset_all (x=0, z=random if randomizing) src/stim/simulators/frame_simulator.inl:152-163
  - measurement record + lookback            src/stim/io/measure_record_batch.inl:48-94
  - DETECTOR / OBSERVABLE_INCLUDE            src/stim/simulators/frame_simulator.inl:221-252
  - PAULI_CHANNEL_1/2 disjoint order         src/stim/simulators/tableau_simulator.h:291-324
  - MPP / MXX / MYY / MZZ / SPP semantics    src/stim/circuit/gate_decomposition.cc:88-170,190-274
  - shot-major bit packing (bit m of shot s at byte m//8, bit m%8)  src/stim/py/numpy.pybind.cc:20-63

Two modes:
  * fixed-fault mode (``faults=...``): no random noise is drawn; only the injected faults fire.
    Frame randomization is off (FlipSimulator(disable_stabilizer_randomization=True),
    src/stim/simulators/frame_simulator.pybind.cc:310-329).
  * sampled mode (``rng=numpy.random.Generator``): every channel is drawn densely per site.

Fault addressing (shared with the product's fixed-fault entry point): a fault is
``(instruction_index, group_index, code)`` where ``instruction_index`` counts instructions of the
REPEAT-unrolled circuit and ``group_index`` is the channel site inside that instruction:
  single-qubit channels (X/Y/Z_ERROR, DEPOLARIZE1, PAULI_CHANNEL_1): target index, code bit0 = X, bit1 = Z
  two-qubit channels (DEPOLARIZE2, PAULI_CHANNEL_2): pair index, code bits (x1, z1, x2, z2)
  noisy results (M/MR/MX/.../MPP/MXX/MPAD with p): result index, code 1 = flip the recorded result
  E / ELSE_CORRELATED_ERROR: group 0, code 1 = the error fires (ELSE only if no earlier link fired)
  HERALDED_ERASE / HERALDED_PAULI_CHANNEL_1: target index, code bit2 = herald, bit0 = X, bit1 = Z
"""

import numpy as np

from . import stim_text as st

ONE_Q_SWAP = {'H', 'H_NXZ', 'SQRT_Y', 'SQRT_Y_DAG'}           # do_H_XZ   frame_simulator.inl:344-350
ONE_Q_ZX = {'S', 'S_DAG', 'H_XY', 'H_NXY'}                    # do_H_XY   :352-358  z ^= x
ONE_Q_XZ = {'SQRT_X', 'SQRT_X_DAG', 'H_YZ', 'H_NYZ'}          # do_H_YZ   :360-366  x ^= z
ONE_Q_CXYZ = {'C_XYZ', 'C_NXYZ', 'C_XNYZ', 'C_XYNZ'}          # do_C_XYZ  :368-375
ONE_Q_CZYX = {'C_ZYX', 'C_NZYX', 'C_ZNYX', 'C_ZYNX'}          # do_C_ZYX  :377-384
NO_EFFECT = {'I', 'X', 'Y', 'Z', 'II', 'I_ERROR', 'II_ERROR', 'TICK', 'QUBIT_COORDS', 'SHIFT_COORDS'}


def _pauli_bits(p):
    return {'X': (1, 0), 'Y': (1, 1), 'Z': (0, 1)}[p]


def product_terms(group):
    """Combines a '*'-joined Pauli product into {qubit: (x, z)} plus classical bits.

    Raises on anti-Hermitian products, as accumulate_next_obs_terms_to_pauli_string_helper does
    (gate_decomposition.cc:39-86)."""
    acc = {}
    bits = []
    imag = 0
    for t in group:
        if t.kind == st.PAULI:
            x, z = _pauli_bits(t.pauli)
            ox, oz = acc.get(t.value, (0, 0))
            if (ox or oz) and (ox, oz) != (x, z):
                imag ^= 1
            acc[t.value] = (ox ^ x, oz ^ z)
        elif t.kind in (st.REC, st.SWEEP):
            bits.append(t)
        else:
            raise ValueError(f'Unsupported product target {t!r}')
    if imag:
        raise ValueError('Anti-Hermitian Pauli product')
    return {q: v for q, v in sorted(acc.items()) if v != (0, 0)}, bits


class FrameOracle:
    def __init__(self, text_or_flat, randomize=False):
        self.flat = st.parse_flat(text_or_flat) if isinstance(text_or_flat, str) else text_or_flat
        self.stats = st.stats(self.flat)
        self.randomize = randomize

    # ------------------------------------------------------------------ helpers
    def _rand_bits(self, rng, n):
        return rng.integers(0, 2, size=n, dtype=np.uint8).astype(bool)

    def _bern(self, rng, p, shape):
        if p <= 0:
            return np.zeros(shape, dtype=bool)
        return rng.random(shape) < p

    # ------------------------------------------------------------------ main loop
    def run(self, shots, faults=None, rng=None):
        """Returns (dets[S, D] bool, obs[S, O] bool).

        faults: optional list (length shots) of lists of (instruction_index, group_index, code).
        rng: numpy Generator for sampled mode (ignored for channels when faults is given)."""
        s = self.stats
        Q, S = s['num_qubits'], shots
        fixed = faults is not None
        if not fixed and rng is None:
            rng = np.random.default_rng()
        fmap = {}
        if fixed:
            for shot, lst in enumerate(faults):
                for (k, j, code) in lst:
                    fmap.setdefault(k, []).append((j, shot, code))
        x = np.zeros((Q, S), dtype=bool)
        z = np.zeros((Q, S), dtype=bool)
        if self.randomize:
            z[:] = rng.integers(0, 2, size=(Q, S)).astype(bool)
        rec = []   # list of bool rows (S,)
        dets = []
        obs = np.zeros((s['num_observables'], S), dtype=bool)
        chain = np.zeros(S, dtype=bool)  # last_correlated_error_occurred

        def flips_for(k, count):
            """Per-result flip rows for a noisy-result instruction."""
            f = np.zeros((count, S), dtype=bool)
            for (j, shot, code) in fmap.get(k, []):
                if code & 1:
                    f[j, shot] ^= True
            return f

        for k, ins in enumerate(self.flat):
            n = ins.name
            T = ins.targets
            a = ins.args
            if n in NO_EFFECT:
                continue
            if n in ONE_Q_SWAP:
                for t in T:
                    q = t.value
                    x[q], z[q] = z[q].copy(), x[q].copy()
            elif n in ONE_Q_ZX:
                for t in T:
                    z[t.value] ^= x[t.value]
            elif n in ONE_Q_XZ:
                for t in T:
                    x[t.value] ^= z[t.value]
            elif n in ONE_Q_CXYZ:
                for t in T:
                    q = t.value
                    x[q] ^= z[q]
                    z[q] ^= x[q]
            elif n in ONE_Q_CZYX:
                for t in T:
                    q = t.value
                    z[q] ^= x[q]
                    x[q] ^= z[q]
            elif n in st.PAIR_UNITARY:
                for i in range(0, len(T), 2):
                    self._pair(n, T[i], T[i + 1], x, z, rec)
            elif n in st.RESET_1Q:
                for t in T:
                    q = t.value
                    # The commuting component is kept, and randomized only when randomizing (XOR with
                    # uniform bits has the distribution of the reference's fresh random bits).
                    r = self._rand_bits(rng, S) if self.randomize else np.zeros(S, dtype=bool)
                    if n == 'R':        # do_RZ :266-274
                        x[q] = False
                        z[q] ^= r
                    elif n == 'RX':     # do_RX :210-219
                        z[q] = False
                        x[q] ^= r
                    else:               # do_RY :255-263
                        z[q] ^= r
                        x[q] = z[q].copy()
            elif n in st.MEASURE_1Q:
                p = a[0] if a else 0.0
                fl = flips_for(k, len(T)) if fixed else self._bern(rng, p, (len(T), S))
                for i, t in enumerate(T):
                    q = t.value
                    r = self._rand_bits(rng, S) if self.randomize else np.zeros(S, dtype=bool)
                    # do_MZ/MX/MY :173-208 and do_MRX/MRY/MRZ :276-317: without randomization the
                    # commuting component of a measured or reset qubit is kept.
                    if n in ('M', 'MR'):
                        rec.append(x[q] ^ fl[i])
                        if n == 'MR':
                            x[q] = False
                        z[q] ^= r
                    elif n in ('MX', 'MRX'):
                        rec.append(z[q] ^ fl[i])
                        if n == 'MRX':
                            z[q] = False
                        x[q] ^= r
                    else:  # MY, MRY
                        rec.append(x[q] ^ z[q] ^ fl[i])
                        if n == 'MRY':
                            z[q] ^= r
                            x[q] = z[q].copy()
                        else:
                            x[q] ^= r
                            z[q] ^= r
            elif n == 'MPAD':
                p = a[0] if a else 0.0
                fl = flips_for(k, len(T)) if fixed else self._bern(rng, p, (len(T), S))
                for i in range(len(T)):
                    rec.append(fl[i].copy())
            elif n in st.PAIR_MEASURE or n == 'MPP':
                p = a[0] if a else 0.0
                if n == 'MPP':
                    groups = st.product_groups(T)
                else:
                    pa = n[1]
                    groups = [[st.Target(st.PAULI, T[i].value, False, pa),
                               st.Target(st.PAULI, T[i + 1].value, False, pa)] for i in range(0, len(T), 2)]
                fl = flips_for(k, len(groups)) if fixed else self._bern(rng, p, (len(groups), S))
                for i, g in enumerate(groups):
                    terms, _ = product_terms(g)
                    v = np.zeros(S, dtype=bool)
                    for q, (px, pz) in terms.items():
                        if px:
                            v ^= z[q]
                        if pz:
                            v ^= x[q]
                    rec.append(v ^ fl[i])
                    if self.randomize and terms:
                        r = self._rand_bits(rng, S)
                        for q, (px, pz) in terms.items():
                            if px:
                                x[q] ^= r
                            if pz:
                                z[q] ^= r
            elif n in ('SPP', 'SPP_DAG'):
                for g in st.product_groups(T):
                    terms, bits = product_terms(g)
                    if not terms:
                        continue
                    v = np.zeros(S, dtype=bool)
                    for q, (px, pz) in terms.items():
                        if px:
                            v ^= z[q]
                        if pz:
                            v ^= x[q]
                    for b in bits:
                        if b.kind == st.REC:
                            v ^= rec[len(rec) - b.value]
                    for q, (px, pz) in terms.items():
                        if px:
                            x[q] ^= v
                        if pz:
                            z[q] ^= v
            elif n == 'DETECTOR':
                d = np.zeros(S, dtype=bool)
                for t in T:
                    d ^= self._lookback(rec, t.value)
                dets.append(d)
            elif n == 'OBSERVABLE_INCLUDE':
                o = int(a[0])
                for t in T:
                    if t.kind == st.REC:
                        obs[o] ^= self._lookback(rec, t.value)
                    elif t.kind == st.PAULI:
                        px, pz = _pauli_bits(t.pauli)
                        if px:
                            obs[o] ^= z[t.value]
                        if pz:
                            obs[o] ^= x[t.value]
                    else:
                        raise ValueError('Unexpected OBSERVABLE_INCLUDE target')
            elif n in ('X_ERROR', 'Y_ERROR', 'Z_ERROR', 'DEPOLARIZE1', 'PAULI_CHANNEL_1'):
                codes = self._channel_codes_1q(k, n, a, len(T), S, fixed, fmap, rng)
                for i, t in enumerate(T):
                    c = codes[i]
                    x[t.value] ^= (c & 1).astype(bool)
                    z[t.value] ^= (c & 2).astype(bool)
            elif n in ('DEPOLARIZE2', 'PAULI_CHANNEL_2'):
                npairs = len(T) // 2
                codes = self._channel_codes_2q(k, n, a, npairs, S, fixed, fmap, rng)
                for i in range(npairs):
                    c = codes[i]
                    q1, q2 = T[2 * i].value, T[2 * i + 1].value
                    x[q1] ^= (c & 1).astype(bool)
                    z[q1] ^= (c & 2).astype(bool)
                    x[q2] ^= (c & 4).astype(bool)
                    z[q2] ^= (c & 8).astype(bool)
            elif n in ('E', 'ELSE_CORRELATED_ERROR'):
                if n == 'E':
                    chain[:] = False
                if fixed:
                    fire = np.zeros(S, dtype=bool)
                    for (j, shot, code) in fmap.get(k, []):
                        if code & 1:
                            fire[shot] = True
                else:
                    fire = self._bern(rng, a[0], S)
                fire &= ~chain
                chain |= fire
                for t in T:
                    px, pz = _pauli_bits(t.pauli)
                    if px:
                        x[t.value] ^= fire
                    if pz:
                        z[t.value] ^= fire
            elif n in ('HERALDED_ERASE', 'HERALDED_PAULI_CHANNEL_1'):
                codes = self._herald_codes(k, n, a, len(T), S, fixed, fmap, rng)
                for i, t in enumerate(T):
                    c = codes[i]
                    x[t.value] ^= (c & 1).astype(bool)
                    z[t.value] ^= (c & 2).astype(bool)
                    rec.append((c & 4).astype(bool))
            else:
                raise NotImplementedError(n)

        D = np.array(dets, dtype=bool).reshape(len(dets), S)
        # The flips of every measurement result, in record order (rows = results), for checks of
        # measurement outputs.
        self.last_record = rec
        return D.T.copy(), obs.T.copy()

    # ------------------------------------------------------------------ pieces
    @staticmethod
    def _lookback(rec, k):
        if k == 0 or k > len(rec):
            raise IndexError('Referred to a measurement record before the beginning of time.')
        return rec[len(rec) - k]

    def _pair(self, n, ta, tb, x, z, rec):
        # Reversed-control gates are the forward rule with the operands swapped (do_XCZ, do_YCZ).
        if n == 'XCZ':
            n, ta, tb = 'CX', tb, ta
        elif n == 'YCZ':
            n, ta, tb = 'CY', tb, ta
        abits = ta.kind in (st.REC, st.SWEEP)
        bbits = tb.kind in (st.REC, st.SWEEP)
        if abits or bbits:
            # Classical control (frame_simulator.inl:139-150, 386-474).
            if n in ('CX', 'CY'):
                if bbits:
                    raise ValueError('Controlled gate had a bit as its target')
                bit = self._ctrl(ta, rec, x.shape[1])
                if n == 'CX':
                    x[tb.value] ^= bit
                else:
                    x[tb.value] ^= bit
                    z[tb.value] ^= bit
                return
            if n == 'CZ':
                if abits and bbits:
                    return
                if abits:
                    z[tb.value] ^= self._ctrl(ta, rec, x.shape[1])
                else:
                    z[ta.value] ^= self._ctrl(tb, rec, x.shape[1])
                return
            raise ValueError(f'{n} cannot take classical bits')
        a, b = ta.value, tb.value
        x1, z1, x2, z2 = x[a].copy(), z[a].copy(), x[b].copy(), z[b].copy()
        if n == 'CX':           # single_cx :386-405
            z1 ^= z2; x2 ^= x1
        elif n == 'CY':         # single_cy :407-428
            z1 ^= x2 ^ z2; z2 ^= x1; x2 ^= x1
        elif n == 'CZ':         # do_ZCZ :448-474
            z1 ^= x2; z2 ^= x1
        elif n == 'SWAP':
            x1, x2 = x2, x1; z1, z2 = z2, z1
        elif n in ('ISWAP', 'ISWAP_DAG'):
            dx = x1 ^ x2
            t1 = z1 ^ dx; t2 = z2 ^ dx
            z1, z2 = t2, t1
            x1, x2 = x2, x1
        elif n == 'CXSWAP':
            z2 ^= z1; z1 ^= z2; x1 ^= x2; x2 ^= x1
        elif n == 'CZSWAP':
            x1, x2 = x2, x1; z1, z2 = z2, z1
            z1 ^= x2; z2 ^= x1
        elif n == 'SWAPCX':
            z1 ^= z2; z2 ^= z1; x2 ^= x1; x1 ^= x2
        elif n in ('SQRT_XX', 'SQRT_XX_DAG'):
            dz = z1 ^ z2; x1 ^= dz; x2 ^= dz
        elif n in ('SQRT_YY', 'SQRT_YY_DAG'):
            d = x1 ^ z1 ^ x2 ^ z2
            x1 ^= d; z1 ^= d; x2 ^= d; z2 ^= d
        elif n in ('SQRT_ZZ', 'SQRT_ZZ_DAG'):
            dx = x1 ^ x2; z1 ^= dx; z2 ^= dx
        elif n == 'XCX':
            x1 ^= z2; x2 ^= z1
        elif n == 'XCY':
            x1 ^= x2 ^ z2; x2 ^= z1; z2 ^= z1
        elif n == 'YCX':
            x2 ^= x1 ^ z1; x1 ^= z2; z1 ^= z2
        elif n == 'YCY':
            y1 = x1 ^ z1; y2 = x2 ^ z2
            x1 ^= y2; z1 ^= y2; x2 ^= y1; z2 ^= y1
        elif n == 'II':
            pass
        else:
            raise NotImplementedError(n)
        x[a], z[a], x[b], z[b] = x1, z1, x2, z2

    def _ctrl(self, t, rec, S):
        if t.kind == st.REC:
            return self._lookback(rec, t.value)
        return np.zeros(S, dtype=bool)  # sweep bits default to False (no sweep table)

    def _channel_codes_1q(self, k, n, a, nt, S, fixed, fmap, rng):
        codes = np.zeros((nt, S), dtype=np.uint8)
        if fixed:
            for (j, shot, code) in fmap.get(k, []):
                codes[j, shot] ^= code & 3
            return codes
        if n == 'X_ERROR':
            codes[self._bern(rng, a[0], (nt, S))] = 1
        elif n == 'Z_ERROR':
            codes[self._bern(rng, a[0], (nt, S))] = 2
        elif n == 'Y_ERROR':
            codes[self._bern(rng, a[0], (nt, S))] = 3
        elif n == 'DEPOLARIZE1':
            hit = self._bern(rng, a[0], (nt, S))
            codes[hit] = rng.integers(1, 4, size=int(hit.sum()))
        elif n == 'PAULI_CHANNEL_1':
            px, py, pz = a
            u = rng.random((nt, S))
            codes[u < px] = 1
            codes[(u >= px) & (u < px + py)] = 3
            codes[(u >= px + py) & (u < px + py + pz)] = 2
        return codes

    def _channel_codes_2q(self, k, n, a, npairs, S, fixed, fmap, rng):
        codes = np.zeros((npairs, S), dtype=np.uint8)
        if fixed:
            for (j, shot, code) in fmap.get(k, []):
                codes[j, shot] ^= code & 15
            return codes
        if n == 'DEPOLARIZE2':
            hit = self._bern(rng, a[0], (npairs, S))
            codes[hit] = rng.integers(1, 16, size=int(hit.sum()))
        else:
            # PAULI_CHANNEL_2 argument order IX, IY, IZ, XI, ... (tableau_simulator.h:291-324):
            # pauli index i in 1..15, first qubit = (i >> 2) & 3, second = i & 3, 1=X 2=Y 3=Z.
            u = rng.random((npairs, S))
            acc = 0.0
            for i in range(1, 16):
                p = a[i - 1]
                m = (u >= acc) & (u < acc + p)
                acc += p
                c1 = {0: 0, 1: 1, 2: 3, 3: 2}[(i >> 2) & 3]
                c2 = {0: 0, 1: 1, 2: 3, 3: 2}[i & 3]
                codes[m] = c1 | (c2 << 2)
        return codes

    def _herald_codes(self, k, n, a, nt, S, fixed, fmap, rng):
        codes = np.zeros((nt, S), dtype=np.uint8)
        if fixed:
            for (j, shot, code) in fmap.get(k, []):
                codes[j, shot] = code & 7
            return codes
        if n == 'HERALDED_ERASE':
            hit = self._bern(rng, a[0], (nt, S))
            codes[hit] = 4 | rng.integers(0, 4, size=int(hit.sum()))
        else:
            pi, px, py, pz = a
            u = rng.random((nt, S))
            codes[u < pi] = 4
            codes[(u >= pi) & (u < pi + px)] = 5
            codes[(u >= pi + px) & (u < pi + px + py)] = 7
            codes[(u >= pi + px + py) & (u < pi + px + py + pz)] = 6
        return codes


def pack_bits(rows):
    """Shot-major bit packing: bit m of shot s -> byte m//8, bit m%8 (numpy.pybind.cc:20-63)."""
    rows = np.asarray(rows, dtype=bool)
    return np.packbits(rows, axis=1, bitorder='little')


def noise_sites(flat, allow_chains=False):
    """Enumerates every (instruction_index, group_index, code) basis fault and its probability weight.

    Returns a list of channels; each channel is (instruction_index, group_index, [(code, prob), ...])
    with disjoint outcomes, used for exact detector firing probabilities."""
    chans = []
    for k, ins in enumerate(flat):
        n, a, T = ins.name, ins.args, ins.targets
        if n in ('X_ERROR', 'Y_ERROR', 'Z_ERROR'):
            c = {'X_ERROR': 1, 'Y_ERROR': 3, 'Z_ERROR': 2}[n]
            for j in range(len(T)):
                chans.append((k, j, [(c, a[0])]))
        elif n == 'DEPOLARIZE1':
            for j in range(len(T)):
                chans.append((k, j, [(c, a[0] / 3) for c in (1, 2, 3)]))
        elif n == 'PAULI_CHANNEL_1':
            for j in range(len(T)):
                chans.append((k, j, [(1, a[0]), (3, a[1]), (2, a[2])]))
        elif n == 'DEPOLARIZE2':
            for j in range(len(T) // 2):
                chans.append((k, j, [(c, a[0] / 15) for c in range(1, 16)]))
        elif n == 'PAULI_CHANNEL_2':
            m = {0: 0, 1: 1, 2: 3, 3: 2}
            for j in range(len(T) // 2):
                chans.append((k, j, [(m[(i >> 2) & 3] | (m[i & 3] << 2), a[i - 1]) for i in range(1, 16)]))
        elif (n in st.MEASURE_1Q or n in st.PAIR_MEASURE or n in ('MPP', 'MPAD')) and a and a[0] > 0:
            for j in range(st.num_results(ins)):
                chans.append((k, j, [(1, a[0])]))
        elif n == 'HERALDED_ERASE':
            for j in range(len(T)):
                chans.append((k, j, [(4 | c, a[0] / 4) for c in range(4)]))
        elif n == 'HERALDED_PAULI_CHANNEL_1':
            for j in range(len(T)):
                chans.append((k, j, [(4, a[0]), (5, a[1]), (7, a[2]), (6, a[3])]))
        elif n == 'E':
            chans.append((k, 0, [(1, a[0])]))
        elif n == 'ELSE_CORRELATED_ERROR':
            if not allow_chains:
                raise NotImplementedError('exact rates for correlated chains are not enumerated')
            chans.append((k, 0, [(1, a[0])]))
    return chans


def _channels(flat):
    """Every channel as a list of disjoint outcomes (instruction_index, group_index, code, probability).

    A correlated chain is one channel: CORRELATED_ERROR starts it and each later
    ELSE_CORRELATED_ERROR fires only when no earlier element of the chain fired, so element i fires with
    p_i * prod_{j<i} (1 - p_j) (frame_simulator.inl:746-776: the `last_correlated_error_occurred` mask,
    cleared only by CORRELATED_ERROR)."""
    out = []
    chain = None   # (outcome list, probability that no element fired yet)
    for (k, j, outs) in noise_sites(flat, allow_chains=True):
        name = flat[k].name
        if name in ('E', 'CORRELATED_ERROR'):
            p = outs[0][1]
            chain = [[(k, 0, 1, p)], 1.0 - p]
            out.append(chain[0])
        elif name == 'ELSE_CORRELATED_ERROR':
            if chain is None:
                # No chain yet: the mask of fired elements starts clear, so the element fires freely.
                chain = [[], 1.0]
                out.append(chain[0])
            p = outs[0][1]
            chain[0].append((k, 0, 1, p * chain[1]))
            chain[1] *= 1.0 - p
        else:
            out.append([(k, j, code, p) for (code, p) in outs])
    return out


def exact_detector_rates(text_or_flat, batch=4096):
    """Exact P(detector fires) and P(observable flips) for independent channels of disjoint outcomes.

    Each outcome's symptom comes from the frame oracle in fixed-fault mode (linearity over GF(2)); a
    channel flips an output bit with q_c = sum of its outcomes' probabilities whose symptom holds the
    bit, and P(odd) = (1 - prod_c (1 - 2 q_c)) / 2 per output bit."""
    orc = FrameOracle(text_or_flat)
    chans = _channels(orc.flat)
    items = [it for ch in chans for it in ch]
    D = orc.stats['num_detectors'] + orc.stats['num_observables']
    syms = np.zeros((len(items), D), dtype=bool)
    for b in range(0, len(items), batch):
        chunk = items[b:b + batch]
        dets, obs = orc.run(len(chunk), faults=[[(k, j, c)] for (k, j, c, _) in chunk])
        syms[b:b + len(chunk)] = np.concatenate([dets, obs], axis=1)
    log_prod = np.zeros(D)
    sign = np.ones(D)
    pos = 0
    for ch in chans:
        q = np.zeros(D)
        for (_, _, _, p) in ch:
            q += p * syms[pos]
            pos += 1
        f = 1 - 2 * q
        sign = np.where(f < 0, -sign, sign)
        with np.errstate(divide='ignore'):
            log_prod += np.log(np.abs(f))
    return (1 - sign * np.exp(log_prod)) / 2
