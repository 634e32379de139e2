"""Detector error models and their sampler, with the reference's Python API.

Reference: stim.DetectorErrorModel (src/stim/dem/detector_error_model.pybind.cc) and
stim.CompiledDemSampler (src/stim/simulators/dem_sampler.pybind.cc:40-309). Every error mechanism of
the model fires independently with its probability and flips its detectors and observables
(src/stim/simulators/dem_sampler.inl:50-77). Sampling runs on the GPU through the C ABI
(mfs_dem_compile); there is no CPU path.
"""

import ctypes
import os

import numpy as np

from . import _native as N
from .sampler import CompiledDetectorSampler as _CDS


def _text_bytes(text):
    return text.encode('utf-8') if isinstance(text, str) else bytes(text)


class DetectorErrorModel:
    def __init__(self, text: str = ''):
        self._text = str(text)
        t = _text_bytes(self._text)
        out = (ctypes.c_uint64 * 3)()
        N.check(N.lib().mfs_dem_parse_stats(t, len(t), out))
        self.num_detectors, self.num_observables, self.num_errors = list(out)

    @staticmethod
    def from_file(path):
        with open(path) as f:
            return DetectorErrorModel(f.read())

    def __str__(self):
        return self._text

    def compile_sampler(self, *, seed=None) -> 'CompiledDemSampler':
        return CompiledDemSampler(self, seed=seed)


class CompiledDemSampler:
    def __init__(self, model, *, seed=None):
        if not isinstance(model, DetectorErrorModel):
            model = DetectorErrorModel(str(model))
        self._model = model
        t = _text_bytes(str(model))
        has_seed = seed is not None
        if has_seed:
            seed = int(seed)
            if seed < 0 or seed >= 1 << 64:
                raise ValueError('seed must be a non-negative 64-bit integer')
        h = ctypes.c_void_p()
        N.check(N.lib().mfs_dem_compile(t, len(t), seed if has_seed else 0, 1 if has_seed else 0, ctypes.byref(h)))
        self._h = h
        self.num_detectors = model.num_detectors
        self.num_observables = model.num_observables
        self.num_errors = model.num_errors

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            N.lib().mfs_destroy(h)
            self._h = None

    # Device-side introspection and timing, shared with the circuit sampler (they read only the handle).
    engine_info = _CDS.engine_info
    bench_device = _CDS.bench_device

    def _replay_faults(self, errs, shots):
        """(shot offsets, fault triples) for a recorded-error array: bool (shots, E) or bit-packed uint8
        (shots, ceil(E/8))."""
        E = self.num_errors
        errs = np.asarray(errs)
        if errs.dtype == np.bool_:
            if errs.shape != (shots, E):
                raise ValueError('recorded_errors_to_replay.shape != (shots, num_errors)')
            bits = errs
        elif errs.dtype == np.uint8:
            if errs.shape != (shots, (E + 7) // 8):
                raise ValueError('recorded_errors_to_replay.shape != (shots, ceil(num_errors / 8))')
            bits = np.unpackbits(errs, axis=1, bitorder='little')[:, :E].astype(bool)
        else:
            raise ValueError('recorded_errors_to_replay must be a bool or uint8 array')
        rows, cols = np.nonzero(bits)
        offsets = np.zeros(shots + 1, dtype=np.uint64)
        np.cumsum(np.bincount(rows, minlength=shots), out=offsets[1:])
        faults = np.zeros((len(cols), 3), dtype=np.uint32)
        faults[:, 0] = cols
        faults[:, 2] = 1
        return np.ascontiguousarray(offsets), np.ascontiguousarray(faults.reshape(-1))

    def sample(self, shots: int, *, bit_packed: bool = False, return_errors: bool = False,
               recorded_errors_to_replay=None):
        """Returns (detection events, observable flips, None): bool (shots, n) arrays, or uint8
        (shots, ceil(n/8)) bit-packed little-endian rows."""
        if return_errors:
            raise ValueError('return_errors=True is not supported by this sampler')
        shots = int(shots)
        if shots < 0:
            raise ValueError('shots must be non-negative')
        D, O = self.num_detectors, self.num_observables
        if bit_packed:
            dets = np.zeros((shots, (D + 7) // 8), dtype=np.uint8)
            obs = np.zeros((shots, (O + 7) // 8), dtype=np.uint8)
        else:
            dets = np.zeros((shots, D), dtype=np.bool_)
            obs = np.zeros((shots, O), dtype=np.bool_)
        flags = N.MFS_SEPARATE_OBSERVABLES | (N.MFS_BIT_PACKED if bit_packed else 0)
        d_stride = max(dets.shape[1], 1)
        o_stride = max(obs.shape[1], 1)
        d_ptr = dets.ctypes.data if dets.size else None
        o_ptr = obs.ctypes.data if obs.size else None
        if shots and (D or O):
            if recorded_errors_to_replay is None:
                N.check(N.lib().mfs_sample(self._h, shots, flags, d_ptr, d_stride, o_ptr, o_stride))
            else:
                offsets, faults = self._replay_faults(recorded_errors_to_replay, shots)
                N.check(N.lib().mfs_sample_with_faults(self._h, shots, offsets.ctypes.data,
                                                       faults.ctypes.data if faults.size else None, flags,
                                                       d_ptr, d_stride, o_ptr, o_stride))
        elif recorded_errors_to_replay is not None:
            self._replay_faults(recorded_errors_to_replay, shots)   # shape checks only: no outputs to fill
        return dets, obs, None

    def sample_write(self, shots: int, *, det_out_file, det_out_format: str = '01', obs_out_file,
                     obs_out_format: str = '01', err_out_file=None, err_out_format: str = '01',
                     replay_err_in_file=None, replay_err_in_format: str = '01'):
        """Samples and writes detection events and observable flips in the reference's result formats."""
        if err_out_file is not None:
            raise ValueError('err_out_file is not supported by this sampler')
        shots = int(shots)
        if replay_err_in_file is None:
            N.check(N.lib().mfs_sample_write(self._h, shots, os.fsencode(det_out_file), det_out_format.encode(), 0, 0,
                                             os.fsencode(obs_out_file), obs_out_format.encode()))
            return
        errs = _read_bits(replay_err_in_file, replay_err_in_format, shots, self.num_errors)
        dets, obs, _ = self.sample(shots, recorded_errors_to_replay=errs)
        _write_bits(det_out_file, det_out_format, dets)
        _write_bits(obs_out_file, obs_out_format, obs)


def _read_bits(path, fmt, shots, n):
    """Reads shots x n bits in the '01' or 'b8' result format (doc/result_formats.md)."""
    data = open(path, 'rb').read()
    if fmt == '01':
        lines = [l for l in data.decode().split('\n') if l != '']
        if len(lines) != shots or any(len(l) != n for l in lines):
            raise ValueError('replay data does not have the expected shape')
        return np.array([[c == '1' for c in l] for l in lines], dtype=np.bool_).reshape(shots, n)
    if fmt == 'b8':
        row = (n + 7) // 8
        if len(data) != shots * row:
            raise ValueError('replay data does not have the expected size')
        return np.frombuffer(data, dtype=np.uint8).reshape(shots, row).copy()
    raise ValueError(f'replay format {fmt!r} is not supported (use 01 or b8)')


def _write_bits(path, fmt, rows):
    if fmt == '01':
        with open(path, 'w') as f:
            for r in rows:
                f.write(''.join('1' if b else '0' for b in r) + '\n')
    elif fmt == 'b8':
        packed = np.packbits(np.asarray(rows, dtype=np.uint8), axis=1, bitorder='little') if len(rows) else b''
        with open(path, 'wb') as f:
            f.write(bytes(packed))
    else:
        raise ValueError(f'format {fmt!r} is not supported with replayed errors (use 01 or b8)')
