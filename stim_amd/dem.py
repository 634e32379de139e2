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
from . import formats
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
        """Returns (detection events, observable flips, errors): bool (shots, n) arrays, or uint8
        (shots, ceil(n/8)) bit-packed little-endian rows. errors is None unless return_errors=True; it
        then holds which mechanisms fired in each shot (the replayed ones when replaying), as
        src/stim/simulators/dem_sampler.pybind.cc:45-79 returns them."""
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
        E = self.num_errors
        errs = None
        if return_errors and recorded_errors_to_replay is None:
            errs = np.zeros((shots, (E + 7) // 8), dtype=np.uint8) if bit_packed else np.zeros((shots, E), np.bool_)
            if shots:
                N.check(N.lib().mfs_dem_sample_errors(self._h, shots, flags, d_ptr, d_stride, o_ptr, o_stride,
                                                      errs.ctypes.data if errs.size else None, max(errs.shape[1], 1)))
        elif shots and (D or O):
            if recorded_errors_to_replay is None:
                N.check(N.lib().mfs_sample(self._h, shots, flags, d_ptr, d_stride, o_ptr, o_stride))
            else:
                offsets, faults = self._replay_faults(recorded_errors_to_replay, shots)
                N.check(N.lib().mfs_sample_with_faults(self._h, shots, offsets.ctypes.data,
                                                       faults.ctypes.data if faults.size else None, flags,
                                                       d_ptr, d_stride, o_ptr, o_stride))
        elif recorded_errors_to_replay is not None:
            self._replay_faults(recorded_errors_to_replay, shots)   # shape checks only: no outputs to fill
        if return_errors and recorded_errors_to_replay is not None:
            errs = self._errors_as(recorded_errors_to_replay, bit_packed)
        return dets, obs, errs

    def _errors_as(self, errs, bit_packed):
        """A (validated) recorded-error array in the requested layout."""
        a = np.asarray(errs)
        if a.dtype == np.bool_:
            bits = a
        else:
            bits = np.unpackbits(a, axis=1, bitorder='little')[:, :self.num_errors].astype(np.bool_)
        return np.packbits(bits, axis=1, bitorder='little') if bit_packed else bits.copy()

    def sample_write(self, shots: int, *, det_out_file, det_out_format: str = '01', obs_out_file,
                     obs_out_format: str = '01', err_out_file=None, err_out_format: str = '01',
                     replay_err_in_file=None, replay_err_in_format: str = '01'):
        """Samples and writes detection events, observable flips and (err_out_file) the fired mechanisms
        in the reference's result formats, or replays errors read from replay_err_in_file
        (src/stim/simulators/dem_sampler.inl:75-128: prefixes D, L and M in the dets format)."""
        shots = int(shots)
        D, O, E = self.num_detectors, self.num_observables, self.num_errors
        if replay_err_in_file is None and err_out_file is None:
            N.check(N.lib().mfs_sample_write(self._h, shots, os.fsencode(det_out_file), det_out_format.encode(), 0, 0,
                                             os.fsencode(obs_out_file), obs_out_format.encode()))
            return
        if replay_err_in_file is not None:
            errs = formats.read_rows(os.fspath(replay_err_in_file), replay_err_in_format, E, [('M', E)])
            if len(errs) < shots:
                raise ValueError('Expected more error data for the requested number of shots.')
            errs = np.ascontiguousarray(errs[:shots])
            dets, obs, _ = self.sample(shots, recorded_errors_to_replay=errs)
        else:
            dets, obs, errs = self.sample(shots, return_errors=True)
        if err_out_file is not None:
            formats.write_rows(os.fspath(err_out_file), err_out_format, errs, [('M', E)])
        if obs_out_file is not None:
            formats.write_rows(os.fspath(obs_out_file), obs_out_format, obs, [('L', O)])
        if det_out_file is not None:
            formats.write_rows(os.fspath(det_out_file), det_out_format, dets, [('D', D)])
