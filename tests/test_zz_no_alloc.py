"""The allocation-free entry point (mfs_sample_noalloc) against mfs_sample, through the C ABI.

mfs_sample_noalloc stages through a fixed device region instead of hipMalloc'd buffers and splits
calls into batches of whole RNG tiles; its rows must equal mfs_sample's bit for bit on samplers with
the same seed at the same stream position, call after call (the second call smaller than the first).
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from stim_amd import _native as N
    L = N.lib()
    L.mfs_sample_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p,
                                     ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
    L.mfs_sample_noalloc.restype = ctypes.c_int
    return N, L


def _circuit(d=7, p=5e-3, code='surface_code:rotated_memory_z'):
    from stim_amd import gen
    return gen.generate(code, distance=d, rounds=d, after_clifford_depolarization=p,
                        after_reset_flip_probability=p, before_measure_flip_probability=p)


def _pair(text, engine, seed=17):
    import stim_amd
    out = []
    for _ in range(2):
        s = stim_amd.Circuit(text).compile_detector_sampler(seed=seed)
        s.set_engine(engine)
        if engine != 'frame':
            s.engine_info()   # builds the symptom table (allocates) before the allocation-free calls
        out.append(s)
    return out


def _run(entry, s, shots, flags, ncols_det, ncols_obs, packed):
    import torch
    N, L = _lib()
    drow = (ncols_det + 7) // 8 if packed else ncols_det
    orow = (ncols_obs + 7) // 8 if packed else ncols_obs
    dets = torch.full((shots, max(drow, 1)), 0xA5, dtype=torch.uint8, device='cuda:0')
    obs = torch.full((shots, max(orow, 1)), 0xA5, dtype=torch.uint8, device='cuda:0')
    fn = L.mfs_sample_noalloc if entry == 'noalloc' else L.mfs_sample
    N.check(fn(s._h, shots, flags | N.MFS_DEVICE_OUTPUT, dets.data_ptr(), dets.stride(0), obs.data_ptr(),
               obs.stride(0)))
    torch.cuda.synchronize()
    return dets.cpu().numpy(), obs.cpu().numpy()


CASES = [
    # (engine, flags name, packed)
    ('symptom', 'separate', True),
    ('symptom', 'separate', False),
    ('symptom', 'append', False),
    ('frame', 'separate', True),
    ('frame', 'separate', False),
    ('frame', 'append', True),      # packed with a non-byte-aligned observable column: staged unpacked
]


@pytest.mark.parametrize('engine,layout,packed', CASES)
def test_noalloc_equals_sample_on_consecutive_calls(engine, layout, packed):
    N, L = _lib()
    # Rotated surface codes have a multiple of 8 detectors; the repetition code d=6 has 35, so appended
    # observables start inside a byte.
    text = _circuit(d=6, code='repetition_code:memory') if layout == 'append' else _circuit()
    a, b = _pair(text, engine)
    D, O = a.num_detectors, a.num_observables
    flags = (N.MFS_BIT_PACKED if packed else 0) | (
        N.MFS_SEPARATE_OBSERVABLES if layout == 'separate' else N.MFS_APPEND_OBSERVABLES)
    ncols = D + (O if layout == 'append' else 0)
    if layout == 'append':
        assert D % 8 != 0
    # Where the call stages rows (symptom engine with bool rows: packed rows before expansion; frame
    # engine with appended observables inside a byte: unpacked rows before packing), the first call
    # stages more than the fixed region holds, so the allocation-free call runs in several batches.
    first = 120000
    if engine == 'symptom' and not packed:
        first = 2000000
    if engine == 'frame' and layout == 'append':
        first = 2500000
    for shots in (first, 10001):
        want = _run('sample', a, shots, flags, ncols, O, packed)
        got = _run('noalloc', b, shots, flags, ncols, O, packed)
        np.testing.assert_array_equal(got[0], want[0], err_msg=f'{engine} {layout} packed={packed} shots={shots}')
        if layout == 'separate':
            np.testing.assert_array_equal(got[1], want[1])
        assert a.stream_position == b.stream_position
    assert want[0].any()


def test_noalloc_measurement_and_dem_samplers():
    import gzip
    import os
    import stim_amd
    N, L = _lib()
    text = _circuit(d=5)
    ms = [stim_amd.Circuit(text).compile_sampler(seed=4) for _ in range(2)]
    ms[1].engine_info()   # builds the expansion path (allocates) before the allocation-free calls
    M = ms[0].num_measurements
    for shots in (70000, 3000):
        want = _run('sample', ms[0], shots, 0, M, 0, False)
        got = _run('noalloc', ms[1], shots, 0, M, 0, False)
        np.testing.assert_array_equal(got[0], want[0])
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    dem = open(os.path.join(root, 'tests', 'golden', 'dems', 'surface_d5_r5.dem')).read()
    ds = [stim_amd.DetectorErrorModel(dem).compile_sampler(seed=6) for _ in range(2)]
    D, O = ds[0].num_detectors, ds[0].num_observables
    for shots in (300000, 5000):
        want = _run('sample', ds[0], shots, N.MFS_SEPARATE_OBSERVABLES, D, O, False)
        got = _run('noalloc', ds[1], shots, N.MFS_SEPARATE_OBSERVABLES, D, O, False)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])


def test_noalloc_errors():
    import stim_amd
    N, L = _lib()
    s = stim_amd.Circuit(_circuit(d=3)).compile_detector_sampler(seed=1)
    s.set_engine('auto')
    host = np.zeros((64, 8), np.uint8)
    # Host rows are refused (the entry point writes device rows only).
    rc = L.mfs_sample_noalloc(s._h, 64, N.MFS_BIT_PACKED, host.ctypes.data, 8, None, 0)
    assert rc == N.MFS_ERR_INVALID_ARGUMENT
    assert b'MFS_DEVICE_OUTPUT' in L.mfs_last_error()
    # A symptom table that is not built yet is not built here (building allocates).
    import torch
    big = stim_amd.Circuit(_circuit(d=5)).compile_detector_sampler(seed=1)
    dets = torch.empty((1 << 16, 16), dtype=torch.uint8, device='cuda:0')
    rc = L.mfs_sample_noalloc(big._h, 1 << 16, N.MFS_BIT_PACKED | N.MFS_DEVICE_OUTPUT, dets.data_ptr(), 16, None, 0)
    assert rc == N.MFS_ERR_INVALID_ARGUMENT
    assert b'symptom table' in L.mfs_last_error()
    big.engine_info()
    N.check(L.mfs_sample_noalloc(big._h, 1 << 16, N.MFS_BIT_PACKED | N.MFS_DEVICE_OUTPUT, dets.data_ptr(), 16,
                                 None, 0))
