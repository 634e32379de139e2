"""Measurement-to-detection-event conversion on the GPU (Circuit.compile_m2d_converter()).

Pinned by: the reference CLI's `stim m2d` outputs (m2d_vectors.json, bit-exact, with and without the
reference sample, up to the d=25 circuit); the conversion oracle (oracle/m2d_oracle.py, pinned on the
same vectors) on larger random data; and, composed with the measurement sampler, the detector
sampler's fixed-fault vectors (bit-exact) and the exact detector rates of the circuit's error model.
"""

import os

import numpy as np
import pytest

from oracle import m2d_oracle as mo

pytestmark = pytest.mark.gpu


def _rows(hexdata, shots, bits):
    raw = np.frombuffer(bytes.fromhex(hexdata), dtype=np.uint8).reshape(shots, -1)
    return np.unpackbits(raw, axis=1, bitorder='little')[:, :bits].astype(bool)


def _converter(text, skip=False):
    import stim_amd
    return stim_amd.Circuit(text).compile_m2d_converter(skip_reference_sample=skip)


def test_matches_reference_conversions(golden):
    n = 0
    for c in golden.load('m2d_vectors.json'):
        text = c['circuit'] or golden.circuit(c['name'])
        conv = _converter(text, c['skip_reference_sample'])
        M, D, O = conv.num_measurements, conv.num_detectors, conv.num_observables
        meas = _rows(c['measurements'], c['shots'], M)
        want = _rows(c['expected'], c['shots'], D + O)
        got = conv.convert(measurements=meas, append_observables=True)
        np.testing.assert_array_equal(got, want, err_msg=c['name'])
        packed_in = np.packbits(meas, axis=1, bitorder='little')
        got_p = conv.convert(measurements=packed_in, append_observables=True, bit_packed=True)
        np.testing.assert_array_equal(got_p, np.packbits(want, axis=1, bitorder='little'), err_msg=c['name'])
        d, o = conv.convert(measurements=meas, separate_observables=True)
        np.testing.assert_array_equal(d, want[:, :D])
        np.testing.assert_array_equal(o, want[:, D:])
        n += 1
    assert n >= 90


@pytest.mark.parametrize('name', ['color_d5_r3', 'surface_x_d5_r3'])
def test_random_data_matches_oracle(golden, name):
    text = golden.circuit(name)
    conv = _converter(text)
    rng = np.random.default_rng(5)
    meas = rng.random((5000, conv.num_measurements)) < 0.3
    import stim_amd
    ref = stim_amd.Circuit(text).reference_sample()
    np.testing.assert_array_equal(conv.convert(measurements=meas, append_observables=True), mo.convert(text, meas, ref))


def test_windowed_kernel_matches_reference_conversions(golden):
    """The windowed kernel (a ring of recent results per tile) forced on every vector circuit it
    applies to, bit-exact against `stim m2d` (the d=25 circuit among them)."""
    n, saw_d25 = 0, False
    for c in golden.load('m2d_vectors.json'):
        text = c['circuit'] or golden.circuit(c['name'])
        conv = _converter(text, c['skip_reference_sample'])
        info = conv.kernel_info()
        if not info['ring_words_appended'] or not info['ring_words_separate']:
            continue
        conv.set_kernel('windowed')
        M, D, O = conv.num_measurements, conv.num_detectors, conv.num_observables
        meas = _rows(c['measurements'], c['shots'], M)
        want = _rows(c['expected'], c['shots'], D + O)
        np.testing.assert_array_equal(conv.convert(measurements=meas, append_observables=True), want,
                                      err_msg=c['name'])
        d, o = conv.convert(measurements=meas, separate_observables=True)
        np.testing.assert_array_equal(d, want[:, :D], err_msg=c['name'])
        np.testing.assert_array_equal(o, want[:, D:], err_msg=c['name'])
        if info['ring_words_appended_and_separate']:
            da, oa = conv.convert(measurements=meas, separate_observables=True, append_observables=True)
            np.testing.assert_array_equal(da, want, err_msg=c['name'])
            np.testing.assert_array_equal(oa, want[:, D:], err_msg=c['name'])
        n += 1
        saw_d25 |= c['name'] == 'surface_d25_r25'
    assert n >= 20 and saw_d25


@pytest.mark.parametrize('name', ['surface_d5_r5', 'rep_d7_r5', 'color_d5_r3', 'surface_d25_r25'])
def test_windowed_kernel_equals_whole_record_kernel(golden, name):
    """Both kernels on the same random rows, every output layout, a ragged last tile."""
    text = golden.circuit(name)
    conv = _converter(text)
    info = conv.kernel_info()
    assert all(info[k] for k in ('ring_words_appended', 'ring_words_separate', 'ring_words_detectors',
                                 'ring_words_appended_and_separate')), info
    assert info['lds_windowed'] <= 8192 * 4 + 32 * 260
    M = conv.num_measurements
    shots = 1000 if M > 4000 else 4099
    rng = np.random.default_rng(len(name))
    packed = np.packbits(rng.random((shots, M)) < 0.3, axis=1, bitorder='little')
    got = {}
    for kernel in ('whole', 'windowed'):
        conv.set_kernel(kernel)
        got[kernel] = (conv.convert(measurements=packed, bit_packed=True, append_observables=True),
                       conv.convert(measurements=packed, bit_packed=True, separate_observables=True),
                       conv.convert(measurements=packed, bit_packed=True, append_observables=False),
                       conv.convert(measurements=packed, bit_packed=True, append_observables=True,
                                    separate_observables=True))
    for a, b in zip(got['whole'], got['windowed']):
        for x, y in zip(a if isinstance(a, tuple) else (a,), b if isinstance(b, tuple) else (b,)):
            np.testing.assert_array_equal(x, y)
    assert got['whole'][0].any()


def test_records_longer_than_lds_use_the_windowed_kernel():
    """d=9, 600 rounds: 48,081 measurements per shot do not fit one tile's LDS, so only the windowed
    kernel converts them; against the oracle, every output layout, a ragged last tile."""
    import stim_amd
    from stim_amd import gen
    text = gen.generate('surface_code:rotated_memory_z', distance=9, rounds=600, after_clifford_depolarization=1e-3,
                        before_measure_flip_probability=1e-3)
    conv = _converter(text)
    M, D = conv.num_measurements, conv.num_detectors
    assert M == 48081 and D == 48000
    info = conv.kernel_info()
    assert info['lds_whole'] > 160 * 1024 and info['ring_words_appended'] and info['ring_words_separate'], info
    with pytest.raises(ValueError, match='whole-record'):
        conv.set_kernel('whole')
    rng = np.random.default_rng(21)
    meas = rng.random((75, M)) < 0.3
    want = mo.convert(text, meas, stim_amd.Circuit(text).reference_sample())
    np.testing.assert_array_equal(conv.convert(measurements=meas, append_observables=True), want)
    d, o = conv.convert(measurements=meas, separate_observables=True)
    np.testing.assert_array_equal(d, want[:, :D])
    np.testing.assert_array_equal(o, want[:, D:])
    # An observable on the first result: appended, its column is emitted when its block is ready;
    # separate, it would be emitted after the last chunk, when that result has left the ring.
    text2 = text + 'OBSERVABLE_INCLUDE(1) rec[-48081]\n'
    conv2 = _converter(text2)
    info2 = conv2.kernel_info()
    assert info2['ring_words_appended'] == info2['ring_words_separate'] == 0 and info2['ring_words_detectors'], info2
    with pytest.raises(ValueError):
        conv2.convert(measurements=meas, separate_observables=True)
    with pytest.raises(ValueError):
        conv2.convert(measurements=meas, append_observables=True)
    want2 = mo.convert(text2, meas, stim_amd.Circuit(text2).reference_sample())
    np.testing.assert_array_equal(conv2.convert(measurements=meas, append_observables=False), want2[:, :D])


def test_windowed_kernel_refused_where_it_does_not_apply(golden):
    # A detector reaching back over more than the largest ring.
    text = 'M 0\n' + 'M 1\n' * 9000 + 'DETECTOR rec[-1] rec[-9001]\n'
    conv = _converter(text)
    assert conv.kernel_info()['ring_words_appended'] == 0
    with pytest.raises(ValueError, match='windowed'):
        conv.set_kernel('windowed')
    conv.set_kernel('auto')
    meas = np.zeros((3, conv.num_measurements), dtype=bool)
    meas[1, 0] = True
    np.testing.assert_array_equal(conv.convert(measurements=meas, append_observables=False)[:, 0], [False, True, False])
    # Sweep bits keep the whole-record kernel.
    sw = _converter('CX sweep[0] 0\nM 0\nDETECTOR rec[-1]\n')
    assert sw.kernel_info()['ring_words_appended'] == 0


def test_device_rows_aligned_and_unaligned(golden):
    import torch
    from stim_amd import _native as N
    text = golden.circuit('surface_d5_r5')
    conv = _converter(text)
    M, D, O = conv.num_measurements, conv.num_detectors, conv.num_observables
    rng = np.random.default_rng(9)
    meas = rng.random((777, M)) < 0.2
    want = conv.convert(measurements=meas, append_observables=True, bit_packed=True)
    packed = np.packbits(meas, axis=1, bitorder='little')
    for stride in ((M + 7) // 8, 16 * ((M + 127) // 128), 16 * ((M + 127) // 128) + 3):
        host = np.zeros((777, stride), np.uint8)
        host[:, :packed.shape[1]] = packed
        dev_in = torch.from_numpy(host).to('cuda:0')
        dev_out = torch.zeros((777, want.shape[1]), dtype=torch.uint8, device='cuda:0')
        N.check(N.lib().mfs_m2d_convert(conv._h, 777, dev_in.data_ptr(), stride,
                                        N.MFS_BIT_PACKED | N.MFS_APPEND_OBSERVABLES | N.MFS_DEVICE_OUTPUT,
                                        dev_out.data_ptr(), want.shape[1], None, 0))
        np.testing.assert_array_equal(dev_out.cpu().numpy(), want, err_msg=str(stride))


def test_fixed_faults_compose_with_measurement_sampler(golden):
    """m2d(measurement rows under fixed faults) == the detector sampler's fixed-fault detection events."""
    import stim_amd
    fixtures = golden.load('fault_vectors.json')
    n = 0
    for f in fixtures[:12]:
        text = f['circuit']
        faults = [[tuple(x) for x in s['faults']] for s in f['sets']]
        c = stim_amd.Circuit(text)
        rows = c.compile_sampler(seed=1).sample_with_faults(faults)
        got = c.compile_m2d_converter().convert(measurements=rows, append_observables=True)
        want = np.array([[b == '1' for b in s['expected']] for s in f['sets']], dtype=bool)
        np.testing.assert_array_equal(got, want, err_msg=f['name'])
        n += len(faults)
    assert n >= 100


def test_sampled_measurements_convert_to_exact_rates(golden):
    import stim_amd
    from scipy.stats import norm
    rates = np.array(golden.load('dem_rates.json')['surface_d3_r3'])
    c = stim_amd.Circuit(golden.circuit('surface_d3_r3'))
    shots = 1 << 20
    rows = c.compile_sampler(seed=77).sample(shots, bit_packed=True)
    events = c.compile_m2d_converter().convert(measurements=rows, append_observables=True, bit_packed=True)
    bits = np.unpackbits(events, axis=1, bitorder='little')[:, :len(rates)]
    counts = bits.sum(axis=0)
    z = (counts - shots * rates) / np.sqrt(shots * rates * (1 - rates))
    assert np.max(np.abs(z)) < abs(norm.ppf(1e-4 / (2 * len(z))))


def test_d25_full_size_conversion_rates(golden):
    """d=25, 25 rounds: 2^16 sampled measurement shots converted on the GPU; per-detector rates against
    the exact rates of the circuit's error model."""
    import gzip
    import stim_amd
    from scipy.stats import norm
    from oracle import dem_oracle as do
    with gzip.open(os.path.join(golden.dir, 'dems', 'surface_d25_r25_folded.dem.gz'), 'rt') as fh:
        rates = do.exact_rates(do.parse(fh.read()))
    c = stim_amd.Circuit(golden.circuit('surface_d25_r25'))
    shots = 1 << 16
    rows = c.compile_sampler(seed=5).sample(shots, bit_packed=True)
    events = c.compile_m2d_converter().convert(measurements=rows, append_observables=True, bit_packed=True)
    assert events.shape == (shots, (15601 + 7) // 8)
    counts = np.zeros(15608, dtype=np.int64)
    for b in range(8):
        counts[b::8] += ((events >> b) & 1).sum(axis=0, dtype=np.int64)
    counts = counts[:15601]
    z = (counts - shots * rates) / np.sqrt(shots * rates * (1 - rates))
    assert np.max(np.abs(z)) < abs(norm.ppf(1e-4 / (2 * len(z)))), np.argsort(-np.abs(z))[:5]


def test_convert_file_formats(golden, tmp_path):
    from stim_amd import formats
    text = golden.circuit('rep_d7_r5')
    conv = _converter(text)
    M, D, O = conv.num_measurements, conv.num_detectors, conv.num_observables
    rng = np.random.default_rng(3)
    meas = rng.random((128, M)) < 0.1
    want = conv.convert(measurements=meas, append_observables=True)
    for fin in formats.FORMATS:
        formats.write_rows(str(tmp_path / ('m.' + fin)), fin, meas)
        for fout in ('01', 'b8', 'dets', 'hits', 'r8', 'ptb64'):
            out = tmp_path / ('d.' + fout)
            conv.convert_file(measurements_filepath=str(tmp_path / ('m.' + fin)), measurements_format=fin,
                              detection_events_filepath=str(out), detection_events_format=fout,
                              append_observables=True)
            back = formats.read_rows(str(out), fout, D + O, [('D', D), ('L', O)])
            np.testing.assert_array_equal(back, want, err_msg=f'{fin}->{fout}')
    # Separate observables file; dets prefixes.
    conv.convert_file(measurements_filepath=str(tmp_path / 'm.01'), detection_events_filepath=str(tmp_path / 'd2'),
                      detection_events_format='dets', obs_out_filepath=str(tmp_path / 'o2'), obs_out_format='dets')
    first = open(tmp_path / 'd2').read().split('\n')[0]
    assert first == 'shot' + ''.join(' D%d' % k for k in np.flatnonzero(want[0, :D]))
    assert open(tmp_path / 'o2').read().split('\n')[0] == 'shot' + ''.join(' L%d' % k for k in np.flatnonzero(want[0, D:]))


def test_argument_errors(golden):
    conv = _converter(golden.circuit('surface_d3_r3'))
    ms = np.zeros((4, conv.num_measurements), dtype=bool)
    with pytest.raises(ValueError, match='explicitly specify either separate'):
        conv.convert(measurements=ms)
    with pytest.raises(ValueError):
        conv.convert(measurements=np.zeros((4, conv.num_measurements + 1), dtype=bool), append_observables=False)
    with pytest.raises(ValueError):
        conv.convert(measurements=ms, sweep_bits=np.ones((4, 1), dtype=bool), append_observables=False)
    empty = _converter('')
    d1 = empty.convert(measurements=np.zeros((50, 0), dtype=bool), append_observables=True)
    d2, d3 = empty.convert(measurements=np.zeros((50, 0), dtype=bool), separate_observables=True)
    assert d1.shape == d2.shape == d3.shape == (50, 0)


def test_columns_without_measurements():
    """Detectors and observables over no measurements (a circuit with no measurements at all) are
    constant zero columns, in every output layout and with a partial last tile."""
    conv = _converter('DETECTOR\nDETECTOR(1, 2)\nOBSERVABLE_INCLUDE(0)\nOBSERVABLE_INCLUDE(2)')
    ms = np.zeros((45, 0), dtype=bool)
    d = conv.convert(measurements=ms, append_observables=True)
    assert d.shape == (45, 5) and not d.any()
    p = conv.convert(measurements=np.zeros((45, 0), dtype=np.uint8), bit_packed=True, append_observables=True)
    assert p.shape == (45, 1) and not p.any()
    d, o = conv.convert(measurements=ms, separate_observables=True)
    assert d.shape == (45, 2) and o.shape == (45, 3) and not d.any() and not o.any()


# ---------------------------------------------------------------- sweep data

def test_sweep_matches_reference_conversions(golden, tmp_path):
    """Circuits with sweep-controlled CX/CY/CZ/XCZ/YCZ (also inside REPEAT, mixed with quantum pairs,
    reaching observables through Pauli targets): bit-exact against the reference CLI's
    `stim m2d --sweep` on random measurement and sweep data, with and without the reference sample,
    through the array API (bool and bit-packed, appended and separate observables) and files."""
    from stim_amd import formats
    n = 0
    for c in golden.load('m2d_sweep_vectors.json'):
        conv = _converter(c['circuit'], c['skip_reference_sample'])
        M, D, O, K = conv.num_measurements, conv.num_detectors, conv.num_observables, conv.num_sweep_bits
        assert K == c['num_sweep_bits'], c['name']
        meas = _rows(c['measurements'], c['shots'], M)
        sweep = _rows(c['sweep'], c['shots'], K)
        want = _rows(c['expected'], c['shots'], D + O)
        got = conv.convert(measurements=meas, sweep_bits=sweep, append_observables=True)
        np.testing.assert_array_equal(got, want, err_msg=c['name'])
        got_p = conv.convert(measurements=np.packbits(meas, axis=1, bitorder='little'),
                             sweep_bits=np.packbits(sweep, axis=1, bitorder='little'), append_observables=True,
                             bit_packed=True)
        np.testing.assert_array_equal(got_p, np.packbits(want, axis=1, bitorder='little'), err_msg=c['name'])
        d, o = conv.convert(measurements=meas, sweep_bits=sweep, separate_observables=True)
        np.testing.assert_array_equal(d, want[:, :D])
        np.testing.assert_array_equal(o, want[:, D:])
        if n % 6 == 0:
            mp, sp, dp = tmp_path / 'm.01', tmp_path / 's.b8', tmp_path / 'd.01'
            formats.write_rows(str(mp), '01', meas)
            formats.write_rows(str(sp), 'b8', sweep)
            conv.convert_file(measurements_filepath=mp, measurements_format='01', sweep_bits_filepath=sp,
                              sweep_bits_format='b8', detection_events_filepath=dp, detection_events_format='01',
                              append_observables=True)
            np.testing.assert_array_equal(formats.read_rows(str(dp), '01', D + O), want)
        n += 1
    assert n >= 40


def test_sweep_data_is_linear_and_checked(golden):
    """Larger random data: conversion is affine in the sweep bits and measurements
    (f(m, s) = f(m, 0) ^ f(0, s) ^ f(0, 0)); sweep rows on the device; shape errors."""
    import torch
    from stim_amd import _native as N
    c = golden.load('m2d_sweep_vectors.json')[5]
    conv = _converter(c['circuit'])
    M, D, O, K = conv.num_measurements, conv.num_detectors, conv.num_observables, conv.num_sweep_bits
    rng = np.random.default_rng(11)
    shots = 3001
    meas = rng.random((shots, M)) < 0.5
    sweep = rng.random((shots, K)) < 0.5
    f = lambda m, s: conv.convert(measurements=m, sweep_bits=s, append_observables=True)  # noqa: E731
    zm, zs = np.zeros_like(meas), np.zeros_like(sweep)
    full = f(meas, sweep)
    np.testing.assert_array_equal(full, f(meas, zs) ^ f(zm, sweep) ^ f(zm, zs))
    np.testing.assert_array_equal(f(meas, zs), conv.convert(measurements=meas, append_observables=True))
    assert (f(zm, sweep) != f(zm, zs)).any()
    # Device rows.
    dm = torch.from_numpy(np.packbits(meas, axis=1, bitorder='little')).cuda()
    ds = torch.from_numpy(np.packbits(sweep, axis=1, bitorder='little')).cuda()
    row = (D + O + 7) // 8
    out = torch.zeros((shots, row), dtype=torch.uint8, device='cuda:0')
    flags = N.MFS_BIT_PACKED | N.MFS_APPEND_OBSERVABLES | N.MFS_DEVICE_OUTPUT
    N.check(N.lib().mfs_m2d_convert_sweep(conv._h, shots, dm.data_ptr(), dm.stride(0), ds.data_ptr(), ds.stride(0),
                                          flags, out.data_ptr(), row, None, 0))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), np.packbits(full, axis=1, bitorder='little'))
    with pytest.raises(ValueError):
        conv.convert(measurements=meas, sweep_bits=sweep[:, :-1], append_observables=True)
    with pytest.raises(ValueError):
        conv.convert(measurements=meas, sweep_bits=sweep[:-1], append_observables=True)
