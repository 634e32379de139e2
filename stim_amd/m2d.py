"""Measurement-to-detection-event conversion on the GPU, with the reference's Python API.

Reference: stim.CompiledMeasurementsToDetectionEventsConverter
(src/stim/simulators/measurements_to_detection_events.pybind.cc:33-166, 210-380) created by
stim.Circuit.compile_m2d_converter (src/stim/circuit/circuit.pybind.cc:781-785). Conversion runs
through the C ABI (mfs_m2d_*), with sweep data (sweep_bits / sweep_bits_filepath) as in the reference.
"""

import ctypes
import os

import numpy as np

from . import _native as N
from . import formats


class CompiledMeasurementsToDetectionEventsConverter:
    def __init__(self, circuit, *, skip_reference_sample: bool = False):
        from .sampler import Circuit, _text_bytes
        if not isinstance(circuit, Circuit):
            circuit = Circuit(str(circuit))
        self._circuit = circuit
        self._skip = bool(skip_reference_sample)
        h = ctypes.c_void_p()
        t = _text_bytes(str(circuit))
        N.check(N.lib().mfs_m2d_compile(t, len(t), 1 if self._skip else 0, ctypes.byref(h)))
        self._h = h
        out = (ctypes.c_uint64 * 4)()
        N.check(N.lib().mfs_m2d_stats(self._h, out))
        self.num_measurements, self.num_detectors, self.num_observables, self.num_sweep_bits = list(out)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            N.lib().mfs_m2d_destroy(h)
            self._h = None

    def __repr__(self):
        extra = ', skip_reference_sample=True' if self._skip else ''
        return f'stim_amd.CompiledMeasurementsToDetectionEventsConverter(stim_amd.Circuit({str(self._circuit)!r}){extra})'

    def _packed_measurements(self, measurements):
        m = np.asarray(measurements)
        M = self.num_measurements
        if m.ndim != 2:
            raise ValueError('measurements must be a 2d numpy array')
        if m.dtype == np.bool_:
            if m.shape[1] != M:
                raise ValueError(f'measurements.shape[1] != num_measurements ({m.shape[1]} != {M})')
            return np.ascontiguousarray(np.packbits(m, axis=1, bitorder='little')) if M else np.zeros((m.shape[0], 1), np.uint8)
        if m.dtype == np.uint8:
            if m.shape[1] != (M + 7) // 8:
                raise ValueError(f'bit packed measurements.shape[1] != ceil(num_measurements / 8)')
            return np.ascontiguousarray(m) if M else np.zeros((m.shape[0], 1), np.uint8)
        raise ValueError('measurements must have dtype np.bool_ or np.uint8 (bit packed)')

    def _packed_sweep_bits(self, sweep_bits, shots):
        """Sweep data as bit-packed rows of ceil(K/8) bytes (K = num_sweep_bits), or None when all zero.
        Accepts bool (shots, K) or bit-packed uint8 (shots, ceil(K/8)) arrays, as the reference does
        (numpy_array_to_transposed_simd_table, src/stim/py/numpy.pybind.cc)."""
        sw = np.asarray(sweep_bits)
        K = self.num_sweep_bits
        if sw.ndim != 2:
            raise ValueError('sweep_bits must be a 2d numpy array')
        if sw.shape[0] != shots:
            raise ValueError('Need sweep_bits.shape[0] == measurements.shape[0]')
        if sw.dtype == np.bool_:
            if sw.shape[1] != K:
                raise ValueError(f'sweep_bits.shape[1] != num_sweep_bits ({sw.shape[1]} != {K})')
            packed = np.packbits(sw, axis=1, bitorder='little')
        elif sw.dtype == np.uint8:
            if sw.shape[1] != (K + 7) // 8:
                raise ValueError('bit packed sweep_bits.shape[1] != ceil(num_sweep_bits / 8)')
            packed = sw
        else:
            raise ValueError('sweep_bits must have dtype np.bool_ or np.uint8 (bit packed)')
        if K == 0:
            return None
        return np.ascontiguousarray(packed)

    def convert(self, *, measurements, sweep_bits=None, separate_observables=None, append_observables=None,
                bit_packed: bool = False, bit_pack_result: bool = False):
        """Detection events (and observable flips) of measurement data: bool (shots, n) or bit-packed uint8
        rows; a tuple (dets, obs) with separate_observables=True."""
        if separate_observables is None and append_observables is None:
            raise ValueError('To ignore observable flip data, you must explicitly specify either '
                             'separate_observables=False or append_observables=False.')
        separate = bool(separate_observables)
        append = bool(append_observables)
        packed = bool(bit_packed or bit_pack_result)
        meas = self._packed_measurements(measurements)
        shots = meas.shape[0]
        sweep = self._packed_sweep_bits(sweep_bits, shots) if sweep_bits is not None else None
        D, O = self.num_detectors, self.num_observables
        n_out = D + (O if append else 0)
        dtype = np.uint8 if packed else np.bool_
        dets = np.zeros((shots, (n_out + 7) // 8 if packed else n_out), dtype=dtype)
        obs = np.zeros((shots, (O + 7) // 8 if packed else O), dtype=dtype) if separate else None
        flags = (N.MFS_BIT_PACKED if packed else 0) | (N.MFS_APPEND_OBSERVABLES if append else 0) | \
                (N.MFS_SEPARATE_OBSERVABLES if separate else 0)
        if shots:
            N.check(N.lib().mfs_m2d_convert_sweep(
                self._h, shots, meas.ctypes.data, meas.shape[1],
                sweep.ctypes.data if sweep is not None else None, sweep.shape[1] if sweep is not None else 0, flags,
                dets.ctypes.data if dets.size else None, max(dets.shape[1], 1),
                obs.ctypes.data if obs is not None and obs.size else None,
                max(obs.shape[1], 1) if obs is not None else 0))
        if separate:
            return dets, obs
        return dets

    def convert_file(self, *, measurements_filepath, measurements_format: str = '01', sweep_bits_filepath=None,
                     sweep_bits_format: str = '01', detection_events_filepath, detection_events_format: str = '01',
                     append_observables: bool = False, obs_out_filepath=None, obs_out_format: str = '01'):
        """Reads measurements from a file, converts them, and writes detection events (and observables)."""
        rows = formats.read_rows(os.fspath(measurements_filepath), measurements_format, self.num_measurements)
        sweep = None
        if sweep_bits_filepath is not None:
            sweep = formats.read_rows(os.fspath(sweep_bits_filepath), sweep_bits_format, self.num_sweep_bits)
            if len(sweep) != len(rows):
                raise ValueError('The sweep data file holds a different number of shots than the measurement file')
        separate = obs_out_filepath is not None
        out = self.convert(measurements=rows, sweep_bits=sweep, separate_observables=separate,
                           append_observables=append_observables)
        dets, obs = out if separate else (out, None)
        D, O = self.num_detectors, self.num_observables
        prefixes = [('D', D)] + ([('L', O)] if append_observables else [])
        formats.write_rows(os.fspath(detection_events_filepath), detection_events_format, dets, prefixes)
        if separate:
            formats.write_rows(os.fspath(obs_out_filepath), obs_out_format, obs, [('L', O)])

    def set_kernel(self, kernel: str) -> None:
        """'auto' (windowed where it applies and needs less than half the LDS), 'whole' (the tile's whole
        record in LDS) or 'windowed' (a ring of recent results; conversions it does not apply to then
        raise). Results are identical; this only selects the GPU kernel."""
        modes = {'auto': 0, 'whole': 1, 'windowed': 2}
        if kernel not in modes:
            raise ValueError(f'kernel must be one of {sorted(modes)}')
        N.check(N.lib().mfs_m2d_set_kernel(self._h, modes[kernel]))

    def kernel_info(self) -> dict:
        """Ring words of the windowed kernel per output layout (0: not applicable) and LDS bytes per tile."""
        out = (ctypes.c_uint64 * 6)()
        N.check(N.lib().mfs_m2d_kernel_info(self._h, out))
        return {'ring_words_appended': out[0], 'ring_words_separate': out[1], 'ring_words_detectors': out[2],
                'ring_words_appended_and_separate': out[3], 'lds_whole': out[4], 'lds_windowed': out[5]}

    def bench_device(self, shots: int, warmup: int = 2, iters: int = 5) -> float:
        """Average kernel time (ms) of converting `shots` device-resident rows (observables appended)."""
        out = (ctypes.c_double * 1)()
        N.check(N.lib().mfs_m2d_bench(self._h, int(shots), int(warmup), int(iters), out))
        return out[0]
