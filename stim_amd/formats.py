"""Host-side readers and writers of the reference's result formats (doc/result_formats.md) for
file-based conversion. Rows are numpy bool arrays of shape (shots, bits_per_shot).

Formats: '01' (one character per bit, newline per shot), 'b8' (bit-packed little-endian bytes per
shot), 'r8' (per shot, run lengths of zeros before each one, 255 continuing a run, a final run to
one past the end), 'hits' (comma-separated indices of ones per line), 'dets' ('shot' followed by
prefixed indices such as ' D3 L0'), 'ptb64' (per group of 64 shots, 8 little-endian bytes per bit
holding that bit for the 64 shots).
"""

import numpy as np

FORMATS = ('01', 'b8', 'r8', 'ptb64', 'hits', 'dets')


def read_rows(path, fmt, bits_per_shot, prefixes=None):
    """Reads every shot of a file. prefixes (dets format): a list of (letter, count) covering the
    row in order, e.g. [('D', num_detectors), ('L', num_observables)]; default [('M', bits_per_shot)]."""
    n = int(bits_per_shot)
    with open(path, 'rb') as f:
        data = f.read()
    if fmt == '01':
        lines = data.decode().split('\n')
        if lines and lines[-1] == '':
            lines.pop()
        out = np.zeros((len(lines), n), dtype=np.bool_)
        for i, line in enumerate(lines):
            if len(line) != n or any(c not in '01' for c in line):
                raise ValueError(f'01 data line {i} does not have {n} zeros and ones')
            out[i] = np.frombuffer(line.encode(), dtype=np.uint8) == ord('1')
        return out
    if fmt == 'b8':
        row = (n + 7) // 8
        if row == 0:
            raise ValueError('b8 data with zero bits per shot has no shot boundaries')
        if len(data) % row:
            raise ValueError('b8 data is not a whole number of shots')
        raw = np.frombuffer(data, dtype=np.uint8).reshape(-1, row)
        return np.unpackbits(raw, axis=1, bitorder='little')[:, :n].astype(np.bool_)
    if fmt == 'r8':
        shots = []
        cur = []
        pos = 0
        for b in data:
            pos += b
            if b == 255:
                continue
            if pos < n:
                cur.append(pos)
                pos += 1
            elif pos == n:
                shots.append(cur)
                cur, pos = [], 0
            else:
                raise ValueError('r8 data runs past the end of a shot')
        if cur or pos:
            raise ValueError('r8 data ends in the middle of a shot')
        out = np.zeros((len(shots), n), dtype=np.bool_)
        for i, s in enumerate(shots):
            out[i, s] = True
        return out
    if fmt == 'hits':
        lines = data.decode().split('\n')
        if lines and lines[-1] == '':
            lines.pop()
        out = np.zeros((len(lines), n), dtype=np.bool_)
        for i, line in enumerate(lines):
            for tok in filter(None, line.split(',')):
                k = int(tok)
                if not 0 <= k < n:
                    raise ValueError(f'hit {k} out of range')
                out[i, k] ^= True
        return out
    if fmt == 'dets':
        prefixes = prefixes or [('M', n)]
        base = {}
        acc = 0
        for letter, count in prefixes:
            base[letter] = (acc, count)
            acc += count
        lines = data.decode().split('\n')
        if lines and lines[-1] == '':
            lines.pop()
        out = np.zeros((len(lines), n), dtype=np.bool_)
        for i, line in enumerate(lines):
            toks = line.split()
            if not toks or toks[0] != 'shot':
                raise ValueError(f'dets line {i} does not start with "shot"')
            for tok in toks[1:]:
                if tok[0] not in base:
                    # measure_record_reader.inl:573-576 rejects an unknown prefix the same way.
                    raise ValueError(f'Unrecognized DETS prefix {tok[0]!r} in {tok!r}; expected one of '
                                     f'{", ".join(base)}')
                start, count = base[tok[0]]
                k = int(tok[1:])
                if not 0 <= k < count:
                    raise ValueError(f'{tok} out of range')
                out[i, start + k] ^= True
        return out
    if fmt == 'ptb64':
        words = np.frombuffer(data, dtype='<u8')
        if n == 0 or len(words) % n:
            raise ValueError('ptb64 data is not a whole number of 64-shot groups')
        groups = words.reshape(-1, n)
        bits = np.unpackbits(groups.view(np.uint8).reshape(len(groups), n, 8), axis=2, bitorder='little')
        return bits.transpose(0, 2, 1).reshape(-1, n).astype(np.bool_)
    raise ValueError(f'Unrecognized format {fmt!r}; expected one of {", ".join(FORMATS)}')


def write_rows(path, fmt, rows, prefixes=None):
    """Writes rows (bool array, shots x bits) in a result format (see read_rows for prefixes)."""
    rows = np.asarray(rows, dtype=np.bool_)
    shots, n = rows.shape
    if fmt == 'b8':
        with open(path, 'wb') as f:
            f.write(np.packbits(rows, axis=1, bitorder='little').tobytes())
        return
    if fmt == 'ptb64':
        if shots % 64:
            raise ValueError('shots must be a multiple of 64 to use ptb64 format.')
        g = rows.reshape(-1, 64, n).transpose(0, 2, 1)
        with open(path, 'wb') as f:
            f.write(np.packbits(g, axis=2, bitorder='little').tobytes())
        return
    with open(path, 'wb') as f:
        if fmt == '01':
            for r in rows:
                f.write(np.where(r, ord('1'), ord('0')).astype(np.uint8).tobytes() + b'\n')
        elif fmt == 'hits':
            for r in rows:
                f.write((','.join(str(k) for k in np.flatnonzero(r)) + '\n').encode())
        elif fmt == 'dets':
            prefixes = prefixes or [('M', n)]
            for r in rows:
                parts = ['shot']
                acc = 0
                for letter, count in prefixes:
                    parts += ['%s%d' % (letter, k) for k in np.flatnonzero(r[acc:acc + count])]
                    acc += count
                f.write((' '.join(parts) + '\n').encode())
        elif fmt == 'r8':
            for r in rows:
                prev = -1
                for k in list(np.flatnonzero(r)) + [n]:
                    run = int(k - prev - 1)
                    while run >= 255:
                        f.write(b'\xff')
                        run -= 255
                    f.write(bytes([run]))
                    prev = k
        else:
            raise ValueError(f'Unrecognized format {fmt!r}; expected one of {", ".join(FORMATS)}')
