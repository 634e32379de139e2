"""Drop-in detection-event sampler with the reference's Python call shapes.

Mirrors stim.Circuit.compile_detector_sampler and stim.CompiledDetectorSampler
(src/stim/circuit/circuit.pybind.cc:827-830; src/stim/py/compiled_detector_sampler.pybind.cc:34-421):
same argument names, shapes, dtypes, bit order and exceptions. All sampling runs on the GPU
through libstim_amd.so; nothing here computes samples on the CPU.
"""

import ctypes
import os
import pathlib

import numpy as np

from . import _native as N


def _text_bytes(text):
    return text.encode('utf-8') if isinstance(text, str) else bytes(text)


class Circuit:
    """A stabilizer circuit held as .stim text (the boundary input format).

    Reference: stim.Circuit (src/stim/circuit/circuit.h:42). Only what the detection-event path
    needs is provided: text round-trip, stats and compile_detector_sampler."""

    def __init__(self, stim_program_text: str = ''):
        self._text = str(stim_program_text)
        self._stats = None

    @staticmethod
    def from_file(path):
        with open(path) as f:
            return Circuit(f.read())

    @staticmethod
    def generated(code_task: str, *, distance: int, rounds: int, after_clifford_depolarization: float = 0.0,
                  before_round_data_depolarization: float = 0.0, before_measure_flip_probability: float = 0.0,
                  after_reset_flip_probability: float = 0.0) -> 'Circuit':
        """Reference: stim.Circuit.generated (src/stim/circuit/circuit.pybind.cc:1337-1383)."""
        from . import gen
        return Circuit(gen.generate(code_task, distance=distance, rounds=rounds,
                                    after_clifford_depolarization=after_clifford_depolarization,
                                    before_round_data_depolarization=before_round_data_depolarization,
                                    before_measure_flip_probability=before_measure_flip_probability,
                                    after_reset_flip_probability=after_reset_flip_probability))

    def __str__(self):
        return self._text

    def __repr__(self):
        return f'stim_amd.Circuit({self._text!r})'

    def __add__(self, other):
        return Circuit(self._text.rstrip('\n') + '\n' + str(other))

    def _compute_stats(self):
        if self._stats is None:
            out = (ctypes.c_uint64 * 10)()
            t = _text_bytes(self._text)
            N.check(N.lib().mfs_parse_stats(t, len(t), out))
            self._stats = list(out)
        return self._stats

    @property
    def num_qubits(self):
        return self._compute_stats()[0]

    @property
    def num_measurements(self):
        return self._compute_stats()[1]

    @property
    def num_detectors(self):
        return self._compute_stats()[2]

    @property
    def num_observables(self):
        return self._compute_stats()[3]

    def compile_detector_sampler(self, *, seed=None) -> 'CompiledDetectorSampler':
        return CompiledDetectorSampler(self, seed=seed)

    def compile_sampler(self, *, skip_reference_sample: bool = False, seed=None,
                        reference_sample=None) -> 'CompiledMeasurementSampler':
        """Reference: stim.Circuit.compile_sampler (src/stim/circuit/circuit.pybind.cc:588-594)."""
        return CompiledMeasurementSampler(self, skip_reference_sample=skip_reference_sample, seed=seed,
                                          reference_sample=reference_sample)

    def compile_m2d_converter(self, *, skip_reference_sample: bool = False):
        """Reference: stim.Circuit.compile_m2d_converter (src/stim/circuit/circuit.pybind.cc:781-785)."""
        from .m2d import CompiledMeasurementsToDetectionEventsConverter
        return CompiledMeasurementsToDetectionEventsConverter(self, skip_reference_sample=skip_reference_sample)

    def reference_sample(self, *, bit_packed: bool = False) -> np.ndarray:
        """The noiseless run with every random measurement result taken as 0 (host tableau).

        Reference: stim.Circuit.reference_sample (src/stim/circuit/circuit.pybind.cc:664-671)."""
        n = self.num_measurements
        packed = np.zeros((n + 7) // 8, dtype=np.uint8)
        t = _text_bytes(self._text)
        N.check(N.lib().mfs_reference_sample(t, len(t), packed.ctypes.data if packed.size else None, packed.size))
        if bit_packed:
            return packed
        return np.unpackbits(packed, bitorder='little')[:n].astype(np.bool_)


def _check_out_buffer(buf, shape, bit_packed):
    """Validates a caller-provided output buffer (numpy.pybind.cc:28-44, 72-88)."""
    if bit_packed:
        if not isinstance(buf, np.ndarray) or buf.dtype != np.uint8:
            raise ValueError("Output buffer wasn't a numpy.ndarray[np.uint8].")
    else:
        if not isinstance(buf, np.ndarray) or buf.dtype != np.bool_:
            raise ValueError("Output buffer wasn't a numpy.ndarray[np.bool_].")
    if buf.ndim != 2:
        raise ValueError("Output buffer wasn't two dimensional.")
    if tuple(buf.shape) != tuple(shape):
        raise ValueError(f'Expected output buffer to have shape=({shape[0]}, {shape[1]}) '
                         f'but its shape is ({buf.shape[0]}, {buf.shape[1]}).')


class CompiledDetectorSampler:
    """An analyzed stabilizer circuit whose detection events can be sampled quickly on the GPU."""

    def __init__(self, circuit, *, seed=None):
        if not isinstance(circuit, Circuit):
            circuit = Circuit(str(circuit))
        self._circuit = circuit
        handle = ctypes.c_void_p()
        t = _text_bytes(str(circuit))
        has_seed = seed is not None
        if has_seed:
            seed = int(seed)
            if seed < 0 or seed >= 1 << 64:
                raise ValueError('seed must be a non-negative 64-bit integer')
        N.check(N.lib().mfs_compile(t, len(t), seed if has_seed else 0, 1 if has_seed else 0, ctypes.byref(handle)))
        self._h = handle
        stats = (ctypes.c_uint64 * 5)()
        N.check(N.lib().mfs_circuit_stats(self._h, stats))
        self.num_qubits, self.num_measurements, self.num_detectors, self.num_observables, _ = list(stats)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            N.lib().mfs_destroy(h)
            self._h = None

    def __repr__(self):
        return f'stim_amd.CompiledDetectorSampler(stim_amd.Circuit({str(self._circuit)!r}))'

    # ------------------------------------------------------------------ sample
    def _run(self, shots, flags, det_shape, obs_shape, dets_out, obs_out, bit_packed, faults=None):
        dtype = np.uint8 if bit_packed else np.bool_
        dets = np.empty(det_shape, dtype=dtype) if dets_out is None else dets_out
        obs = None
        if obs_shape is not None:
            obs = np.empty(obs_shape, dtype=dtype) if obs_out is None else obs_out
        d_work = dets if dets.flags.c_contiguous else np.empty(det_shape, dtype=dtype)
        o_work = None
        if obs is not None:
            o_work = obs if obs.flags.c_contiguous else np.empty(obs_shape, dtype=dtype)
        d_ptr = d_work.ctypes.data if d_work.size else None
        o_ptr = o_work.ctypes.data if (o_work is not None and o_work.size) else None
        d_stride = det_shape[1] if det_shape[1] else 1
        o_stride = (obs_shape[1] if obs_shape[1] else 1) if obs_shape is not None else 0
        if faults is None:
            N.check(N.lib().mfs_sample(self._h, shots, flags, d_ptr, d_stride, o_ptr, o_stride))
        else:
            offsets, flat = faults
            N.check(N.lib().mfs_sample_with_faults(self._h, shots, offsets.ctypes.data,
                                                   flat.ctypes.data if flat.size else None, flags,
                                                   d_ptr, d_stride, o_ptr, o_stride))
        if d_work is not dets:
            dets[...] = d_work
        if obs is not None and o_work is not obs:
            obs[...] = o_work
        return dets, obs

    def _shapes(self, shots, prepend, append, separate, bit_packed, want_obs):
        D, O = self.num_detectors, self.num_observables
        n = D + O * (int(prepend) + int(append))
        det_shape = (shots, (n + 7) // 8 if bit_packed else n)
        obs_shape = (shots, (O + 7) // 8 if bit_packed else O) if want_obs else None
        return det_shape, obs_shape

    def sample(self, shots: int, *, prepend_observables: bool = False, append_observables: bool = False,
               separate_observables: bool = False, bit_packed: bool = False, dets_out=None, obs_out=None):
        """Returns detection events; see CompiledDetectorSampler.sample in the reference
        (compiled_detector_sampler.pybind.cc:216-322)."""
        return self._sample_impl(shots, prepend_observables, append_observables, separate_observables, bit_packed,
                                 dets_out, obs_out, None)

    def _sample_impl(self, shots, prepend, append, separate, bit_packed, dets_out, obs_out, faults):
        shots = int(shots)
        if shots < 0:
            raise ValueError('shots must be non-negative')
        if separate and (append or prepend):
            raise ValueError(
                "Can't specify separate_observables=True with append_observables=True or prepend_observables=True")
        want_obs = separate or obs_out is not None
        det_shape, obs_shape = self._shapes(shots, prepend, append, separate, bit_packed, want_obs)
        if dets_out is not None:
            _check_out_buffer(dets_out, det_shape, bit_packed)
        if obs_out is not None:
            _check_out_buffer(obs_out, obs_shape, bit_packed)
        flags = (N.MFS_BIT_PACKED if bit_packed else 0) | (N.MFS_PREPEND_OBSERVABLES if prepend else 0) | \
                (N.MFS_APPEND_OBSERVABLES if append else 0)
        if want_obs and not (prepend or append):
            flags |= N.MFS_SEPARATE_OBSERVABLES
        dets, obs = self._run(shots, flags, det_shape, obs_shape if (want_obs and not (prepend or append)) else None,
                              dets_out, obs_out if not (prepend or append) else None, bit_packed, faults)
        if want_obs and (prepend or append):
            # obs_out given together with prepend/append: fill it with a separate-observables call
            # would draw new randomness; instead extract the observable columns of these shots.
            D, O = self.num_detectors, self.num_observables
            cols = np.unpackbits(dets, axis=1, bitorder='little') if bit_packed else dets
            oc = cols[:, 0:O] if prepend else cols[:, D:D + O]
            obs = np.packbits(oc.astype(bool), axis=1, bitorder='little') if bit_packed else oc.astype(np.bool_)
            if obs_out is not None:
                obs_out[...] = obs
                obs = obs_out
        if separate:
            return dets, obs
        return dets

    def sample_bit_packed(self, shots: int, *, prepend_observables: bool = False, append_observables: bool = False):
        """Legacy alias (compiled_detector_sampler.pybind.cc:324-355)."""
        return self.sample(shots, prepend_observables=prepend_observables, append_observables=append_observables,
                           bit_packed=True)

    def sample_write(self, shots: int, *, filepath, format: str = '01', prepend_observables: bool = False,
                     append_observables: bool = False, obs_out_filepath=None, obs_out_format: str = '01'):
        """Samples and writes to a file (compiled_detector_sampler.pybind.cc:88-141, 357-415)."""
        if isinstance(filepath, pathlib.PurePath):
            filepath = str(filepath)
        if isinstance(obs_out_filepath, pathlib.PurePath):
            obs_out_filepath = str(obs_out_filepath)
        if not isinstance(filepath, str):
            raise ValueError(f"Don't know how to write to {filepath!r}")
        if obs_out_filepath is not None and not isinstance(obs_out_filepath, str):
            raise ValueError(f"Don't know how to write observables to {obs_out_filepath!r}")
        N.check(N.lib().mfs_sample_write(self._h, int(shots), os.fsencode(filepath), format.encode(),
                                         int(bool(prepend_observables)), int(bool(append_observables)),
                                         os.fsencode(obs_out_filepath) if obs_out_filepath else None,
                                         obs_out_format.encode()))

    # ------------------------------------------------------------------ fixed-fault mode
    def sample_with_faults(self, faults, *, prepend_observables=False, append_observables=False,
                           separate_observables=False, bit_packed=False):
        """Detection events with exactly the given faults per shot and no other noise.

        faults: a list (one entry per shot) of lists of (instruction_index, site_index, code), where
        instruction_index counts instructions of the REPEAT-unrolled circuit. No reference
        counterpart (the closest is FlipSimulator.broadcast_pauli_errors with stabilizer
        randomization disabled)."""
        shots = len(faults)
        offsets = np.zeros(shots + 1, dtype=np.uint64)
        flat = []
        for i, lst in enumerate(faults):
            for f in lst:
                flat.extend(int(v) for v in f)
            offsets[i + 1] = len(flat) // 3
        flat = np.asarray(flat, dtype=np.uint32)
        return self._sample_impl(shots, prepend_observables, append_observables, separate_observables, bit_packed,
                                 None, None, (offsets, flat))

    # ------------------------------------------------------------------ tuning / introspection
    def set_tile_words(self, words: int) -> int:
        rc = N.lib().mfs_set_tile_words(self._h, int(words))
        if rc < 0:
            N.check(-rc)
        return rc

    def info(self):
        out = (ctypes.c_uint64 * 8)()
        N.check(N.lib().mfs_get_info(self._h, out))
        keys = ['tile_words', 'lds_bytes', 'ops', 'syncs', 'noise_ops', 'randomize', 'ring', 'dense_qubits']
        return dict(zip(keys, list(out)))

    def set_engine(self, engine: str):
        """'auto', 'frame' or 'symptom' (see include/stim_amd.h mfs_set_engine)."""
        N.check(N.lib().mfs_set_engine(self._h, {'auto': 0, 'frame': 1, 'symptom': 2}[engine]))

    def set_symptom_tile_kb(self, kb: int):
        N.check(N.lib().mfs_set_symptom_tile_kb(self._h, int(kb)))

    def set_scan_variant(self, variant: str):
        """'auto' (specialised symptom scan kernel where it applies) or 'generic' (rows are identical)."""
        N.check(N.lib().mfs_set_scan_variant(self._h, {'auto': 0, 'generic': 1}[variant]))

    @property
    def stream_position(self) -> int:
        """Global shot-tile index of the random stream where the next call starts."""
        v = ctypes.c_uint64()
        N.check(N.lib().mfs_get_stream_position(self._h, ctypes.byref(v)))
        return v.value

    @stream_position.setter
    def stream_position(self, tiles: int):
        N.check(N.lib().mfs_set_stream_position(self._h, int(tiles)))

    def engine_info(self):
        out = (ctypes.c_uint64 * 10)()
        N.check(N.lib().mfs_engine_info(self._h, out))
        v = list(out)
        return {'engine': {1: 'frame', 2: 'symptom'}[v[0]], 'classes': v[1], 'columns': v[2], 'entries': v[3],
                'build_ms': v[4] / 1000.0, 'sites': v[5], 'tile_shots': v[6], 'lds_bytes_per_tile': v[7],
                'outcome_blocks': v[8], 'outcome_records': v[9]}

    def bench_device(self, shots: int, warmup: int = 2, iters: int = 5, split: bool = False):
        """Average device time (ms) of one sample of `shots` into HBM-resident packed outputs.

        With split=True returns (whole step, output-writing kernel)."""
        out = (ctypes.c_double * 2)()
        N.check(N.lib().mfs_bench_device(self._h, int(shots), int(warmup), int(iters), out))
        return (out[0], out[1]) if split else out[0]


class CompiledMeasurementSampler(CompiledDetectorSampler):
    """Samples every measurement result of a circuit on the GPU: the circuit's reference sample XOR
    the frame simulator's flips, one row per shot in record order.

    Reference: stim.CompiledMeasurementSampler (src/stim/py/compiled_measurement_sampler.pybind.cc:27-80,
    85-150) and sample_batch_measurements (src/stim/simulators/frame_simulator_util.inl:290-372)."""

    def __init__(self, circuit, *, skip_reference_sample: bool = False, seed=None, reference_sample=None):
        if not isinstance(circuit, Circuit):
            circuit = Circuit(str(circuit))
        self._circuit = circuit
        self._skip = bool(skip_reference_sample)
        if reference_sample is not None and self._skip:
            raise ValueError('skip_reference_sample = True but reference_sample is not None.')
        ref = None
        if reference_sample is not None:
            ref = np.asarray(reference_sample)
            m = circuit.num_measurements
            if ref.dtype == np.bool_:
                ref = np.packbits(ref.reshape(-1), bitorder='little')
            elif ref.dtype != np.uint8:
                raise ValueError('reference_sample must be a bool or bit-packed uint8 numpy array')
            ref = np.ascontiguousarray(ref.reshape(-1))
            if ref.size < (m + 7) // 8:
                raise ValueError('reference_sample is shorter than the number of measurements')
        handle = ctypes.c_void_p()
        t = _text_bytes(str(circuit))
        has_seed = seed is not None
        if has_seed:
            seed = int(seed)
            if seed < 0 or seed >= 1 << 64:
                raise ValueError('seed must be a non-negative 64-bit integer')
        N.check(N.lib().mfs_compile_measurements(t, len(t), seed if has_seed else 0, 1 if has_seed else 0,
                                                 1 if self._skip else 0,
                                                 ref.ctypes.data if ref is not None and ref.size else None,
                                                 0 if ref is None else ref.size, ctypes.byref(handle)))
        self._h = handle
        stats = (ctypes.c_uint64 * 5)()
        N.check(N.lib().mfs_circuit_stats(self._h, stats))
        # Output rows are the measurement results (no observables).
        self.num_qubits, self.num_measurements, self.num_detectors, self.num_observables, _ = list(stats)

    def __repr__(self):
        extra = ', skip_reference_sample=True' if self._skip else ''
        return f'stim_amd.CompiledMeasurementSampler(stim_amd.Circuit({str(self._circuit)!r}){extra})'

    def sample(self, shots: int, *, bit_packed: bool = False) -> np.ndarray:
        """(shots, num_measurements) bool, or (shots, ceil(num_measurements / 8)) uint8 bit-packed."""
        shots = int(shots)
        if shots < 0:
            raise ValueError('shots must be non-negative')
        return self._sample_impl(shots, False, False, False, bit_packed, None, None, None)

    def sample_write(self, shots: int, *, filepath, format: str = '01'):
        """Writes results in a reference result format ('01', 'b8', 'r8', 'ptb64', 'hits', 'dets')."""
        N.check(N.lib().mfs_sample_write(self._h, int(shots), os.fsencode(filepath), format.encode(), 0, 0,
                                         None, None))

    def set_engine(self, engine: str):
        """'auto', 'frame' or 'expansion' (detection events from the symptom engine expanded into
        records; see include/stim_amd.h mfs_set_engine)."""
        N.check(N.lib().mfs_set_engine(self._h, {'auto': 0, 'frame': 1, 'expansion': 2}[engine]))

    def engine_info(self):
        out = (ctypes.c_uint64 * 10)()
        N.check(N.lib().mfs_engine_info(self._h, out))
        v = list(out)
        if v[0] != 3:
            return {'engine': 'frame'}
        return {'engine': 'expansion', 'classes': v[1], 'columns': v[2], 'entries': v[3], 'build_ms': v[4] / 1000.0,
                'sites': v[5], 'free_results': v[6], 'phases': v[7], 'items': v[8]}

    def sample_with_faults(self, faults, *, bit_packed=False):
        """Measurement results with exactly the given faults per shot and no other noise (the frame
        simulator's randomization of random results still applies). Same fault addressing as
        CompiledDetectorSampler.sample_with_faults; no reference counterpart."""
        return super().sample_with_faults(faults, bit_packed=bit_packed)
