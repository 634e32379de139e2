"""ctypes binding of libstim_amd.so (the C ABI declared in include/stim_amd.h).

The library is built in-tree by stim_amd/csrc/build.sh. There is no fallback: if the library is
missing, importing the sampler raises ImportError, and sampling without a GPU raises RuntimeError.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libstim_amd.so')

MFS_OK = 0
MFS_ERR_INVALID_ARGUMENT = 1
MFS_ERR_OUT_OF_RANGE = 2
MFS_ERR_RUNTIME = 3

MFS_BIT_PACKED = 1 << 0
MFS_PREPEND_OBSERVABLES = 1 << 1
MFS_APPEND_OBSERVABLES = 1 << 2
MFS_SEPARATE_OBSERVABLES = 1 << 3
MFS_DEVICE_OUTPUT = 1 << 4

EXPORTED = [
    'mfs_compile', 'mfs_destroy', 'mfs_circuit_stats', 'mfs_parse_stats', 'mfs_sample', 'mfs_sample_write',
    'mfs_sample_with_faults', 'mfs_bench_device', 'mfs_set_tile_words', 'mfs_get_info', 'mfs_last_error',
    'mfs_set_engine', 'mfs_engine_info', 'mfs_set_symptom_tile_kb',
    'mfs_batch_create', 'mfs_batch_destroy', 'mfs_batch_add', 'mfs_batch_set_tile_kb', 'mfs_batch_run',
    'mfs_batch_time', 'mfs_batch_autotune', 'mfs_batch_info', 'mfs_batch_set_schedule',
    'mfs_server_create', 'mfs_server_destroy', 'mfs_server_submit', 'mfs_server_step', 'mfs_server_poll',
    'mfs_server_result', 'mfs_server_copy_result', 'mfs_server_release', 'mfs_server_stats', 'mfs_server_replay',
    'mfs_dem_compile', 'mfs_dem_parse_stats', 'mfs_reference_sample', 'mfs_compile_measurements',
    'mfs_m2d_compile', 'mfs_m2d_destroy', 'mfs_m2d_stats', 'mfs_m2d_convert', 'mfs_m2d_bench',
    'mfs_set_stream_position', 'mfs_get_stream_position', 'mfs_rng_counter', 'mfs_philox', 'mfs_build_id',
    'mfs_sample_noalloc', 'mfs_m2d_convert_sweep', 'mfs_m2d_set_kernel', 'mfs_m2d_kernel_info', 'mfs_dem_sample_errors', 'mfs_set_scan_variant',
    'mfs_expand_plan_stats', 'mfs_expand_shots_host',
]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'{LIB_PATH} is not built; run stim_amd/csrc/build.sh (or __graft_entry__.build())')
    # PyTorch ships its own HIP runtime. When torch is installed, load it first so that this library
    # binds to the same runtime (its libamdhip64.so.7 dependency then resolves to the one already
    # loaded); a second runtime loaded before torch's leaves torch unable to see the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    u64, u32, sz, p = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_size_t, ctypes.c_void_p
    L.mfs_compile.argtypes = [ctypes.c_char_p, sz, u64, ctypes.c_int, ctypes.POINTER(p)]
    L.mfs_compile.restype = ctypes.c_int
    L.mfs_destroy.argtypes = [p]
    L.mfs_destroy.restype = None
    L.mfs_circuit_stats.argtypes = [p, ctypes.POINTER(u64)]
    L.mfs_parse_stats.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(u64)]
    L.mfs_reference_sample.argtypes = [ctypes.c_char_p, sz, p, sz]
    L.mfs_expand_plan_stats.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(u64)]
    L.mfs_expand_shots_host.argtypes = [ctypes.c_char_p, sz, u64, p, sz, p, sz, p, sz]
    L.mfs_m2d_compile.argtypes = [ctypes.c_char_p, sz, ctypes.c_int, ctypes.POINTER(p)]
    L.mfs_m2d_destroy.argtypes = [p]
    L.mfs_m2d_destroy.restype = None
    L.mfs_m2d_stats.argtypes = [p, ctypes.POINTER(u64)]
    L.mfs_m2d_convert.argtypes = [p, u64, p, sz, u32, p, sz, p, sz]
    L.mfs_m2d_convert_sweep.argtypes = [p, u64, p, sz, p, sz, u32, p, sz, p, sz]
    L.mfs_m2d_bench.argtypes = [p, u64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.mfs_m2d_set_kernel.argtypes = [p, ctypes.c_int]
    L.mfs_m2d_kernel_info.argtypes = [p, ctypes.POINTER(u64)]
    L.mfs_compile_measurements.argtypes = [ctypes.c_char_p, sz, u64, ctypes.c_int, ctypes.c_int, p, sz,
                                           ctypes.POINTER(p)]
    L.mfs_sample.argtypes = [p, u64, u32, p, sz, p, sz]
    L.mfs_sample_noalloc.argtypes = [p, u64, u32, p, sz, p, sz]
    L.mfs_sample_write.argtypes = [p, u64, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_char_p, ctypes.c_char_p]
    L.mfs_sample_with_faults.argtypes = [p, u64, p, p, u32, p, sz, p, sz]
    L.mfs_bench_device.argtypes = [p, u64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.mfs_set_tile_words.argtypes = [p, ctypes.c_int]
    L.mfs_get_info.argtypes = [p, ctypes.POINTER(u64)]
    L.mfs_set_engine.argtypes = [p, ctypes.c_int]
    L.mfs_engine_info.argtypes = [p, ctypes.POINTER(u64)]
    L.mfs_set_symptom_tile_kb.argtypes = [p, ctypes.c_int]
    L.mfs_batch_create.argtypes = [ctypes.POINTER(p)]
    L.mfs_batch_destroy.argtypes = [p]
    L.mfs_batch_destroy.restype = None
    L.mfs_batch_add.argtypes = [p, p, u64, p, sz, p, sz]
    L.mfs_batch_set_tile_kb.argtypes = [p, ctypes.c_int]
    L.mfs_batch_run.argtypes = [p, ctypes.c_int]
    L.mfs_batch_time.argtypes = [p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.mfs_batch_autotune.argtypes = [p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    L.mfs_batch_info.argtypes = [p, ctypes.POINTER(u64)]
    L.mfs_batch_set_schedule.argtypes = [p, ctypes.c_int]
    dbl = ctypes.c_double
    L.mfs_server_create.argtypes = [u64, dbl, ctypes.POINTER(p)]
    L.mfs_server_destroy.argtypes = [p]
    L.mfs_server_destroy.restype = None
    L.mfs_server_submit.argtypes = [p, p, u64, u64, ctypes.POINTER(u64)]
    L.mfs_server_step.argtypes = [p, ctypes.POINTER(ctypes.c_int)]
    L.mfs_server_poll.argtypes = [p, ctypes.POINTER(u64), u32, ctypes.POINTER(u32)]
    L.mfs_server_result.argtypes = [p, u64, ctypes.POINTER(p), ctypes.POINTER(sz), ctypes.POINTER(p),
                                    ctypes.POINTER(sz), ctypes.POINTER(u64), ctypes.POINTER(dbl)]
    L.mfs_server_copy_result.argtypes = [p, u64, p, sz, p, sz]
    L.mfs_server_release.argtypes = [p, u64]
    L.mfs_server_stats.argtypes = [p, ctypes.POINTER(u64)]
    L.mfs_server_replay.argtypes = [p, u64, ctypes.POINTER(p), ctypes.POINTER(u64), ctypes.POINTER(u64),
                                    ctypes.POINTER(dbl), ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
    L.mfs_dem_compile.argtypes = [ctypes.c_char_p, sz, u64, ctypes.c_int, ctypes.POINTER(p)]
    L.mfs_dem_parse_stats.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(u64)]
    L.mfs_dem_sample_errors.argtypes = [p, u64, u32, p, sz, p, sz, p, sz]
    L.mfs_set_stream_position.argtypes = [p, u64]
    L.mfs_set_scan_variant.argtypes = [p, ctypes.c_int]
    L.mfs_get_stream_position.argtypes = [p, ctypes.POINTER(u64)]
    L.mfs_rng_counter.argtypes = [ctypes.c_int, u64, u64, u32, u32, u32, ctypes.POINTER(u32)]
    L.mfs_philox.argtypes = [p, u64, u64, p]
    L.mfs_build_id.argtypes = []
    L.mfs_build_id.restype = ctypes.c_char_p
    L.mfs_last_error.argtypes = []
    L.mfs_last_error.restype = ctypes.c_char_p
    for name in EXPORTED:
        if name not in ('mfs_destroy', 'mfs_last_error', 'mfs_compile', 'mfs_batch_destroy', 'mfs_server_destroy',
                        'mfs_build_id'):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def build_id():
    """Source hash the loaded library was built from."""
    return lib().mfs_build_id().decode()


def check_build_id():
    """Raises RuntimeError unless the loaded library was built from the sources in this tree."""
    from stim_amd._buildid import source_hash
    want, have = source_hash(), build_id()
    if want != have:
        raise RuntimeError(f'{LIB_PATH} was built from other sources (library {have}, tree {want}); rebuild it')
    return have


def check(rc):
    """Raises the Python exception the reference raises for the same failure."""
    if rc == MFS_OK:
        return
    msg = lib().mfs_last_error().decode('utf-8', 'replace')
    if rc == MFS_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if rc == MFS_ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    raise RuntimeError(msg)
