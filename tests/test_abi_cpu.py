"""Host-side checks of the C ABI library that need no GPU: symbols, parser, stats, errors."""

import ctypes
import os
import re

import pytest

from oracle import stim_text as st

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from stim_amd import _native
    return _native.lib()


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, 'include', 'stim_amd.h')).read()
    declared = set(re.findall(r'\b(mfs_[a-z_0-9]+)\s*\(', header))
    assert len(declared) >= 10
    L = _lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    from stim_amd import _native
    assert declared == set(_native.EXPORTED)


def _stats(text):
    out = (ctypes.c_uint64 * 10)()
    b = text.encode()
    from stim_amd import _native
    _native.check(_lib().mfs_parse_stats(b, len(b), out))
    return list(out)


def test_stats_match_oracle_on_fixtures(golden):
    for f in golden.load('fault_vectors.json'):
        s = _stats(f['circuit'])
        o = f['stats']
        assert s[:5] == [o['num_qubits'], o['num_measurements'], o['num_detectors'], o['num_observables'],
                         o['max_lookback']], f['name']


def test_d25_program_fits_lds(golden):
    s = _stats(golden.circuit('surface_d25_r25'))
    assert s[2] == 15600 and s[3] == 1
    assert s[9] <= 64 * 1024  # 64-shot tile fits several workgroups per CU


@pytest.mark.parametrize('text,exc', [
    ('H 0\nCX 0\n', ValueError),
    ('NOT_A_GATE 0\n', ValueError),
    ('X_ERROR(1.5) 0\n', ValueError),
    ('REPEAT 0 {\nH 0\n}\n', ValueError),
    ('H 0\n}\n', ValueError),
    ('REPEAT 2 {\nH 0\n', ValueError),
    ('MPP X0*Z0\n', ValueError),
    ('M 0\nDETECTOR rec[-2]\n', IndexError),
    ('CX 0 rec[-1]\n', ValueError),
    ('DEPOLARIZE2(0.1) 0 0\n', ValueError),
    ('PAULI_CHANNEL_1(0.5, 0.5, 0.5) 0\n', ValueError),
    ('DETECTOR 0\n', ValueError),
    # frame_simulator_util.test.cc:46-48 (bad_detector): a bare record target is not an instruction.
    ('rec[-1]\n', ValueError),
    # frame_simulator.test.cc:573-643 (quantum_cannot_control_classical): a record target may only be
    # the control of a classically controlled Pauli.
    ('M 0\nCNOT 1 rec[-1]\n', ValueError),
    ('M 0\nCY 1 rec[-1]\n', ValueError),
    ('M 0\nYCZ rec[-1] 1\n', ValueError),
    ('M 0\nXCZ rec[-1] 1\n', ValueError),
    ('M 0\nSWAP 1 rec[-1]\n', ValueError),
])
def test_parse_errors_map_to_reference_exceptions(text, exc):
    with pytest.raises(exc):
        _stats(text)


def test_parser_accepts_reference_syntax():
    text = '''
    # comment
    QUBIT_COORDS(0, 1) 0
    h 0   # lowercase name
    CNOT 0 1 ZCX 1 2
    X_ERROR[tag with spaces](0.1) 0
    MPP !X0*Z1 Y2
    REPEAT 3 {
        REPEAT 2 {
            M 0
        }
        DETECTOR(1, 2) rec[-1] rec[-2]
    }
    OBSERVABLE_INCLUDE(2) rec[-1] X1
    SPP X0*!Z1 Y2
    SPP_DAG Z0
    MXX 0 1
    MYY 1 2
    MZZ 0 2
    HERALDED_ERASE(0.1) 3
    E(0.1) X0 Y1
    ELSE_CORRELATED_ERROR(0.2) Z2
    '''.replace('CNOT 0 1 ZCX 1 2', 'CNOT 0 1\nZCX 1 2')
    s = _stats(text)
    o = st.stats(st.parse_flat(text))
    assert s[:5] == [o['num_qubits'], o['num_measurements'], o['num_detectors'], o['num_observables'],
                     o['max_lookback']]
    assert s[2] == 3 and s[3] == 3


def test_barriers_fewer_than_ops(golden):
    s = _stats(golden.circuit('surface_d5_r5'))
    assert 0 < s[6] < s[5]


def test_huge_repeat_counts_are_bounded():
    """REPEAT blocks are unrolled for the device program, so their size is checked before unrolling:
    empty bodies cost nothing whatever the count, oversized circuits and detector error models fail at
    once with a clear error (instead of growing until memory runs out), and blocks without error
    mechanisms still declare the detectors of their last pass."""
    import ctypes
    import time
    from stim_amd import _native as N
    L = N.lib()

    def run(fn, text, n):
        out = (ctypes.c_uint64 * n)()
        b = text.encode()
        t0 = time.time()
        rc = fn(b, len(b), out)
        assert time.time() - t0 < 5
        return rc, list(out), (L.mfs_last_error().decode() if rc else '')

    rc, out, _ = run(L.mfs_parse_stats, 'REPEAT 1000000000000000000 {\n}\nH 0\n', 10)
    assert rc == 0 and out[0] == 1
    rc, out, err = run(L.mfs_parse_stats, 'REPEAT 1000000000000 {\nM 0\n}\n', 10)
    assert rc == N.MFS_ERR_INVALID_ARGUMENT and 'unroll' in err
    rc, out, err = run(L.mfs_dem_parse_stats, 'repeat 366666666666666666 {\nerror(0.1) D0\nshift_detectors 1\n}\n', 3)
    assert rc == N.MFS_ERR_INVALID_ARGUMENT and 'too large' in err
    rc, out, _ = run(L.mfs_dem_parse_stats, 'repeat 1000000000000000000 {\n}\nerror(0.1) D0\n', 3)
    assert rc == 0 and out == [1, 0, 1]
    # The last pass declares D2 at offset 12; the error after the block is D0 at offset 15.
    rc, out, _ = run(L.mfs_dem_parse_stats, 'repeat 5 {\ndetector D2\nshift_detectors 3\n}\nerror(0.1) D0\n', 3)
    assert rc == 0 and out == [16, 0, 1]
    rc, out, _ = run(L.mfs_dem_parse_stats, 'repeat 5 {\ndetector D2\nshift_detectors 3\n}\n', 3)
    assert rc == 0 and out == [15, 0, 0]
