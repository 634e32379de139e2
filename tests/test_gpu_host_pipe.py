"""Rows returned in host memory travel through a double-buffered pipeline: device slots, pinned
buffers on a copy stream, host threads into the caller's rows. A call that takes several pipeline
batches must return exactly the rows of the same call into device memory (one batch, no pipeline), in
every layout, on both engines, with any host row stride; the detector-error-model sampler's fired
mechanism rows likewise.

Batches hold about 64 MiB of host rows, so the d=25 r=25 circuit (1,951 packed bytes per shot, 15,601
bool bytes) takes several batches at 10^5 packed or 10^4 bool shots.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ENGINES = ['frame', 'symptom']


def _twins(text, engine, seed):
    import stim_amd
    out = []
    for _ in range(2):
        s = stim_amd.Circuit(text).compile_detector_sampler(seed=seed)
        s.set_engine(engine)
        out.append(s)
    return out


def _device_rows(s, shots, flags, det_row, obs_row):
    import torch
    from stim_amd import _native as N
    d = torch.empty((shots, max(det_row, 1)), dtype=torch.uint8, device='cuda')
    o = torch.empty((shots, max(obs_row, 1)), dtype=torch.uint8, device='cuda')
    N.check(N.lib().mfs_sample(s._h, shots, flags | N.MFS_DEVICE_OUTPUT, d.data_ptr(), d.stride(0), o.data_ptr(),
                               o.stride(0)))
    torch.cuda.synchronize()
    return d.cpu().numpy()[:, :det_row], o.cpu().numpy()[:, :obs_row]


@pytest.mark.parametrize('engine', ENGINES)
def test_packed_separate_multi_batch_equals_device_rows(golden, engine):
    from stim_amd import _native as N
    text = golden.circuit('surface_d25_r25')
    host_s, dev_s = _twins(text, engine, 77)
    shots = 100_003
    # Two calls in a row: the second starts where the first's RNG tiles ended on both twins.
    for call in range(2):
        dets, obs = host_s.sample(shots, bit_packed=True, separate_observables=True)
        dd, do = _device_rows(dev_s, shots, N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES, dets.shape[1],
                              obs.shape[1])
        assert np.array_equal(dets, dd), (engine, call)
        assert np.array_equal(obs, do), (engine, call)
        assert dets.any()


@pytest.mark.parametrize('engine', ENGINES)
def test_bool_append_multi_batch_equals_device_rows(golden, engine):
    from stim_amd import _native as N
    text = golden.circuit('surface_d25_r25')
    host_s, dev_s = _twins(text, engine, 78)
    shots = 10_007
    rows = host_s.sample(shots, append_observables=True)
    dd, _ = _device_rows(dev_s, shots, N.MFS_APPEND_OBSERVABLES, rows.shape[1], 0)
    assert rows.dtype == np.bool_
    assert np.array_equal(rows.view(np.uint8), dd)


@pytest.mark.parametrize('engine', ENGINES)
def test_strided_host_rows_multi_batch(golden, engine):
    """Host rows with padding between them (stride above the row's bytes) through the C ABI."""
    from stim_amd import _native as N
    text = golden.circuit('surface_d25_r25')
    host_s, dev_s = _twins(text, engine, 79)
    shots = 70_001
    D, O = host_s.num_detectors, host_s.num_observables
    drow, orow = (D + 7) // 8, (O + 7) // 8
    dbig = np.full((shots, drow + 13), 0xA5, dtype=np.uint8)
    obig = np.full((shots, orow + 3), 0x5A, dtype=np.uint8)
    flags = N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES
    N.check(N.lib().mfs_sample(host_s._h, shots, flags, dbig.ctypes.data, dbig.shape[1], obig.ctypes.data,
                               obig.shape[1]))
    dd, do = _device_rows(dev_s, shots, flags, drow, orow)
    assert np.array_equal(dbig[:, :drow], dd)
    assert np.array_equal(obig[:, :orow], do)
    assert (dbig[:, drow:] == 0xA5).all() and (obig[:, orow:] == 0x5A).all()


def test_dem_errors_multi_batch_equal_device_rows(golden):
    import gzip
    import os
    import torch
    import stim_amd
    from stim_amd import _native as N
    with gzip.open(os.path.join(golden.dir, 'dems', 'surface_d25_r25_folded.dem.gz'), 'rt') as f:
        dem_text = f.read()
    a = stim_amd.DetectorErrorModel(dem_text).compile_sampler(seed=5)
    b = stim_amd.DetectorErrorModel(dem_text).compile_sampler(seed=5)
    shots = 12_001
    dets, obs, errs = a.sample(shots, bit_packed=True, return_errors=True)
    D, O, E = dets.shape[1], obs.shape[1], errs.shape[1]
    dd = torch.empty((shots, D), dtype=torch.uint8, device='cuda')
    do = torch.empty((shots, O), dtype=torch.uint8, device='cuda')
    de = torch.empty((shots, E), dtype=torch.uint8, device='cuda')
    flags = N.MFS_BIT_PACKED | N.MFS_SEPARATE_OBSERVABLES | N.MFS_DEVICE_OUTPUT
    N.check(N.lib().mfs_dem_sample_errors(b._h, shots, flags, dd.data_ptr(), D, do.data_ptr(), O, de.data_ptr(), E))
    torch.cuda.synchronize()
    assert np.array_equal(dets, dd.cpu().numpy())
    assert np.array_equal(obs, do.cpu().numpy())
    assert np.array_equal(errs, de.cpu().numpy())
    assert errs.any()
