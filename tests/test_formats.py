"""Result-format readers and writers of stim_amd.formats (file-based conversion) against the reference
CLI's bytes for every format (formats.json: a deterministic circuit sampled by `stim detect` in each
format, with observables plain, appended, prepended and in a separate file). No GPU is used.
"""

import numpy as np
import pytest

from oracle import stim_text as st
from stim_amd import formats


def _layout(golden):
    f = golden.load('formats.json')
    s = st.stats(st.parse_flat(f['circuit']))
    return f, s['num_detectors'], s['num_observables']


def _rows01(data, n):
    lines = data.decode().split('\n')[:-1]
    return np.array([[c == '1' for c in l] for l in lines], dtype=bool).reshape(len(lines), n)


@pytest.mark.parametrize('fmt', list(formats.FORMATS))
def test_writers_and_readers_match_reference_bytes(golden, tmp_path, fmt):
    f, D, O = _layout(golden)
    checked = 0
    for key, hexdata in f['outputs'].items():
        kfmt, mode, shots = key.split('/')
        if kfmt != fmt:
            continue
        want = bytes.fromhex(hexdata)
        # The reference CLI's dets format always carries the observables, first.
        eff = 'prepend' if (fmt == 'dets' and mode == 'plain') else mode
        n, prefixes = {'plain': (D, [('D', D)]), 'append': (D + O, [('D', D), ('L', O)]),
                       'prepend': (D + O, [('L', O), ('D', D)]), 'obs_out_dets': (D, [('D', D)]),
                       'obs_out_obs': (O, [('L', O)])}[eff]
        # The circuit is deterministic: every shot is the same row.
        row = _rows01(bytes.fromhex(f['outputs']['01/%s/5' % eff]), n)[0]
        rows01 = np.broadcast_to(row, (int(shots), n)).copy()
        path = tmp_path / ('x.' + fmt)
        formats.write_rows(str(path), fmt, rows01, prefixes)
        assert path.read_bytes() == want, key
        path.write_bytes(want)
        np.testing.assert_array_equal(formats.read_rows(str(path), fmt, n, prefixes), rows01, err_msg=key)
        checked += 1
    assert checked >= 3


def test_readers_reject_malformed_data(tmp_path):
    p = tmp_path / 'bad'
    p.write_bytes(b'0102\n')
    with pytest.raises(ValueError):
        formats.read_rows(str(p), '01', 4)
    p.write_bytes(b'\x00\x01\x02')
    with pytest.raises(ValueError):
        formats.read_rows(str(p), 'b8', 9)
    p.write_bytes(b'\x09')
    with pytest.raises(ValueError):
        formats.read_rows(str(p), 'r8', 4)
    with pytest.raises(ValueError):
        formats.read_rows(str(p), 'xyz', 4)


def test_malformed_files_raise_value_error(tmp_path):
    """Seeded byte mutations of valid files of every format either read or raise ValueError (the
    reference's readers throw std::invalid_argument), and never take long."""
    import random
    import time
    from stim_amd import formats as F
    rng = random.Random(7)
    n = 13
    rows = np.random.default_rng(0).integers(0, 2, size=(64, n)).astype(bool)
    for fmt in ['01', 'b8', 'r8', 'hits', 'dets', 'ptb64']:
        p = tmp_path / ('x.' + fmt)
        F.write_rows(str(p), fmt, rows)
        good = p.read_bytes()
        for _ in range(400):
            d = bytearray(good)
            for _ in range(rng.randint(1, 4)):
                if not d:
                    break
                i = rng.randrange(len(d))
                op = rng.randrange(3)
                if op == 0:
                    d[i] = rng.randrange(256)
                elif op == 1:
                    d[i:i] = bytes(rng.choice(b'0123456789 ,\nDLMshot\xff') for _ in range(rng.randint(1, 12)))
                else:
                    del d[i:i + rng.randint(1, 8)]
            p.write_bytes(bytes(d))
            t0 = time.time()
            try:
                F.read_rows(str(p), fmt, n)
            except ValueError:
                pass
            assert time.time() - t0 < 2, fmt
