"""ctypes declarations of libsupertrack_amd's C ABI (include/supertrack_amd.h).

The library is loaded from this package directory only; a missing library is
an error (there is no fallback implementation).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SUPERTRACK_AMD_LIBRARY: another build of the same library (e.g. with host
# sanitizers); it is loaded the same way and there is still no fallback
LIB_PATH = os.environ.get("SUPERTRACK_AMD_LIBRARY") or os.path.join(HERE, "libsupertrack_amd.so")

ST_OK = 0
ST_HIST_LIN = 0
ST_HIST_LOG = 1
ST_F64 = 0
ST_F32 = 1

c_double_p = C.POINTER(C.c_double)
c_int64_p = C.POINTER(C.c_int64)
c_uint64_p = C.POINTER(C.c_uint64)
c_int32_p = C.POINTER(C.c_int32)


class Geometry(C.Structure):
    _fields_ = [("half_length", C.c_double), ("length", C.c_double), ("diameter", C.c_double),
                ("radius", C.c_double), ("gso", C.c_double), ("n_linear", C.c_int32), ("pad", C.c_int32)]


class Track(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("edep", C.c_void_p),
                ("n", C.c_int64), ("capacity", C.c_int64)]


class Pairs(C.Structure):
    _fields_ = [("volume", C.c_void_p), ("edep", C.c_void_p), ("num_elements", C.c_void_p),
                ("capacity", C.c_int64), ("key_bound", C.c_uint64), ("count_dev", C.c_void_p)]


class LibraryInfo(C.Structure):
    _fields_ = [("n_tracks", C.c_int64), ("n_points", C.c_int64), ("device_bytes", C.c_int64),
                ("fast_tracks", C.c_int64), ("exact_tracks", C.c_int64), ("precision", C.c_int32),
                ("n_geometries", C.c_int32)]


class BatchOpts(C.Structure):
    _fields_ = [("n_oversamples", C.c_int32), ("shift_xy", C.c_int32), ("shift_z", C.c_int32),
                ("track_index_base", C.c_uint32), ("seed", C.c_uint64), ("use_graph", C.c_int32),
                ("autotune", C.c_int32), ("block_threads", C.c_int32), ("table_slots", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("points_per_thread", C.c_int32), ("hist_copies", C.c_int32)]


class SynthSpec(C.Structure):
    _fields_ = [("voxel_half_length", C.c_double), ("entry_half_length", C.c_double),
                ("length_min_nm", C.c_double), ("length_max_nm", C.c_double),
                ("spacing_nm", C.c_double), ("p_cluster", C.c_double), ("p_cluster_max", C.c_double),
                ("cluster_mean", C.c_double), ("walk_sigma_nm", C.c_double),
                ("edep_median_ev", C.c_double), ("edep_sigma_ln", C.c_double),
                ("edep_min_ev", C.c_double), ("edep_max_ev", C.c_double),
                ("spacing_min_nm", C.c_double), ("spacing_max_nm", C.c_double),
                ("n_min", C.c_int64), ("n_max", C.c_int64), ("stream_id", C.c_uint32), ("pad", C.c_uint32)]


# name -> (restype, argtypes)
PROTOTYPES = {
    "st_last_error": (C.c_char_p, []),
    "st_version": (C.c_char_p, []),
    "st_device_count": (C.c_int, [c_int32_p]),
    "st_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "st_free": (C.c_int, [C.c_void_p]),
    "st_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "st_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "st_memcpy_dtod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "st_memset": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    "st_device_synchronize": (C.c_int, []),
    "st_default_stream": (C.c_int, [C.POINTER(C.c_void_p)]),
    "st_shutdown": (C.c_int, []),
    "st_debug_fail_allocation": (C.c_int, [C.c_int64]),
    "st_geometry_make": (C.c_int, [C.c_double, C.c_double, C.POINTER(Geometry)]),
    "st_track_alloc": (C.c_int, [C.POINTER(Track), C.c_int64]),
    "st_track_upload": (C.c_int, [C.POINTER(Track), c_double_p, c_double_p, c_double_p, c_double_p, C.c_int64]),
    "st_track_free": (None, [C.POINTER(Track)]),
    "st_pairs_alloc": (C.c_int, [C.POINTER(Pairs), C.c_int64]),
    "st_pairs_free": (None, [C.POINTER(Pairs)]),
    "st_pairs_set_count": (C.c_int, [C.POINTER(Pairs), C.c_int32]),
    "st_vcs_create": (C.c_int, [C.POINTER(Geometry), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                C.POINTER(C.c_void_p)]),
    "st_vcs_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "st_vcs_begin_track": (C.c_int, [C.c_void_p, C.POINTER(Track), C.c_uint64, C.c_uint32, C.c_int32]),
    "st_vcs_process_oversample": (C.c_int, [C.c_void_p, C.POINTER(Track), C.POINTER(Pairs)]),
    "st_vcs_end_track": (C.c_int, [C.c_void_p]),
    "st_vcs_destroy": (None, [C.c_void_p]),
    "st_vcs_set_queued": (C.c_int, [C.c_void_p, C.c_int32]),
    "st_hist_create": (C.c_int, [C.c_int32, C.c_float, C.c_float, C.c_int32, C.POINTER(C.c_void_p)]),
    "st_hist_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "st_hist_nbins": (C.c_int, [C.c_void_p, c_int32_p]),
    "st_hist_edges": (C.c_int, [C.c_void_p, c_double_p]),
    "st_hist_begin_track": (C.c_int, [C.c_void_p, C.POINTER(Pairs)]),
    "st_hist_add": (C.c_int, [C.c_void_p, C.POINTER(Pairs)]),
    "st_hist_end_track": (C.c_int, [C.c_void_p, c_double_p]),
    "st_hist_fetch": (C.c_int, [C.c_void_p, c_uint64_p, C.c_int32]),
    "st_hist_destroy": (None, [C.c_void_p]),
    "st_library_create": (C.c_int, [c_double_p, c_double_p, c_double_p, c_double_p, c_int64_p, C.c_int64,
                                    C.POINTER(Geometry), C.c_int32, c_int32_p, C.c_int32,
                                    C.POINTER(C.c_void_p)]),
    "st_library_begin": (C.c_int, [C.c_int64, C.c_int64, C.POINTER(Geometry), C.c_int32, C.c_int32,
                                   C.POINTER(C.c_void_p)]),
    "st_library_append": (C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p, c_int64_p,
                                    C.c_int64, c_int32_p]),
    "st_library_finish": (C.c_int, [C.c_void_p]),
    "st_library_info_get": (C.c_int, [C.c_void_p, C.POINTER(LibraryInfo)]),
    "st_library_destroy": (None, [C.c_void_p]),
    "st_batch_create": (C.c_int, [C.c_void_p, C.POINTER(BatchOpts), C.POINTER(C.c_void_p), C.c_int32,
                                  C.POINTER(C.c_void_p)]),
    "st_batch_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "st_batch_run": (C.c_int, [C.c_void_p]),
    "st_batch_sync": (C.c_int, [C.c_void_p]),
    "st_batch_params": (C.c_int, [C.c_void_p, c_int32_p]),
    "st_batch_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "st_batch_stats": (C.c_int, [C.c_void_p, c_int64_p]),
    "st_batch_balance": (C.c_int, [C.c_void_p, c_uint64_p]),
    "st_batch_run_static": (C.c_int, [C.c_void_p, C.POINTER(BatchOpts), C.POINTER(C.c_void_p), C.c_int32,
                                      C.c_void_p, c_int64_p]),
    "st_batch_record": (C.c_int, [C.c_void_p, C.c_int64]),
    "st_batch_record_fetch": (C.c_int, [C.c_void_p, c_int64_p, c_uint64_p, c_double_p, c_int64_p]),
    "st_batch_destroy": (None, [C.c_void_p]),
    "st_synth_counts": (C.c_int, [C.POINTER(SynthSpec), C.c_int64, C.c_int64, c_int64_p]),
    "st_synth_fill": (C.c_int, [C.POINTER(SynthSpec), C.c_int64, C.c_int64, c_int64_p, c_double_p, c_double_p,
                                c_double_p, c_double_p]),
}

_lib = None


class SuperTrackError(RuntimeError):
    def __init__(self, code: int, func: str, msg: str):
        super().__init__(f"{func} failed with status {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Loads libsupertrack_amd.so and declares its prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m supertrack_amd.build` "
                          "(there is no non-HIP fallback)")
    handle = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(handle, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return _lib


def check(rc: int, func: str) -> None:
    if rc != ST_OK:
        msg = lib().st_last_error()
        raise SuperTrackError(rc, func, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    fn = getattr(lib(), name)
    check(fn(*args), name)
