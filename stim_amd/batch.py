"""Batched sampling: many compiled samplers (different circuits, different shot counts) sampled by one
persistent kernel, optionally replayed as a HIP graph, with an autotuned tile size.

No reference counterpart: the reference samples one circuit per call
(src/stim/py/compiled_detector_sampler.pybind.cc:34-86). Each job here yields exactly what
`sampler.sample(shots, bit_packed=True, separate_observables=True)` would at the same point of the
sampler's random stream, with the same tile size; the outputs stay in device memory (torch uint8
tensors of shape (shots, ceil(D/8)) and (shots, ceil(O/8))).

    batch = BatchSampler([(s1, 10_000), (s2, 2_000_000)])
    batch.autotune()
    outputs = batch.run(use_graph=True)
"""

import ctypes

from . import _native as N


class BatchSampler:
    def __init__(self, jobs, device='cuda:0', keep_observables=True):
        import torch
        self._torch = torch
        self._h = None
        h = ctypes.c_void_p()
        N.check(N.lib().mfs_batch_create(ctypes.byref(h)))
        self._h = h
        self.jobs = []
        self.outputs = []
        for sampler, shots in jobs:
            shots = int(shots)
            drow = (sampler.num_detectors + 7) // 8
            orow = (sampler.num_observables + 7) // 8
            dets = torch.empty((shots, max(drow, 1)), dtype=torch.uint8, device=device)
            obs = torch.empty((shots, max(orow, 1)), dtype=torch.uint8, device=device) if keep_observables else None
            N.check(N.lib().mfs_batch_add(self._h, sampler._h, shots, dets.data_ptr(), dets.stride(0),
                                          obs.data_ptr() if obs is not None else None,
                                          obs.stride(0) if obs is not None else 0))
            self.jobs.append((sampler, shots))
            self.outputs.append((dets[:, :drow], obs[:, :orow] if obs is not None else None))

    def __del__(self):
        if getattr(self, '_h', None):
            N.lib().mfs_batch_destroy(self._h)
            self._h = None

    def set_tile_kb(self, kb: int):
        N.check(N.lib().mfs_batch_set_tile_kb(self._h, int(kb)))

    def run(self, use_graph: bool = False):
        """Samples every job once; returns [(dets, obs)] device tensors (reused across runs)."""
        N.check(N.lib().mfs_batch_run(self._h, 1 if use_graph else 0))
        return self.outputs

    def time(self, use_graph: bool = False, warmup: int = 1, iters: int = 3):
        """(device ms per run from HIP events, wall ms per run)."""
        out = (ctypes.c_double * 2)()
        N.check(N.lib().mfs_batch_time(self._h, 1 if use_graph else 0, int(warmup), int(iters), out))
        return out[0], out[1]

    def autotune(self, candidates=(12, 16, 20, 24, 32, 48), iters: int = 2):
        """Picks the fastest LDS output-image size per workgroup (KB) and schedule (persistent or one
        workgroup per tile); returns (kb, ms per run)."""
        arr = (ctypes.c_int * len(candidates))(*[int(c) for c in candidates])
        kb, ms = ctypes.c_int(), ctypes.c_double()
        N.check(N.lib().mfs_batch_autotune(self._h, arr, len(candidates), int(iters), ctypes.byref(kb),
                                           ctypes.byref(ms)))
        return kb.value, ms.value

    def set_schedule(self, persistent: bool):
        """Persistent workgroups claiming tiles (True) or one workgroup per tile (False)."""
        N.check(N.lib().mfs_batch_set_schedule(self._h, 1 if persistent else 0))

    def info(self):
        out = (ctypes.c_uint64 * 8)()
        N.check(N.lib().mfs_batch_info(self._h, out))
        keys = ['jobs', 'kernel_jobs', 'separate_jobs', 'tiles', 'workgroups', 'lds_bytes', 'persistent', 'tile_kb']
        return dict(zip(keys, list(out)))
