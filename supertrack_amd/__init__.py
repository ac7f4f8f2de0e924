"""supertrack_amd -- MI355X-native VoxelConstrainedSphere superposition path.

Python view of the C ABI in include/supertrack_amd.h.  Class and method names
follow the reference's C++ interface (SuperTrack: SimulationMethod /
VoxelConstrainedSphereMethod, Histogram, Track, VolumeEdepPair,
SphericalGeometry), so a per-track loop reads like SuperTrackManager.cu:58-83:

    method.allocate_track_process(track, seed, track_index, n_os)
    histogram.allocate_track_process(pairs)
    for _ in range(n_os):
        method.process_track(track, pairs)
        histogram.sort_reduce_and_add_to_histogram(pairs)
    method.free_track_process(); histogram.free_track_process()

All compute runs in libsupertrack_amd.so (HIP, gfx950).  There is no CPU
fallback: without the library every entry point raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _capi
from ._capi import ST_F32, ST_F64, ST_HIST_LIN, ST_HIST_LOG, SuperTrackError, call, lib

__all__ = [
    "SphericalGeometry", "DeviceArray", "Track", "VolumeEdepPair", "VoxelConstrainedSphereMethod",
    "Histogram", "TrackLibrary", "BatchRun", "synth_library", "run_plain", "ST_F32", "ST_F64",
    "ST_HIST_LIN", "ST_HIST_LOG", "SuperTrackError", "device_count", "synchronize", "default_stream_handle",
]


# Wrappers that own library objects, freed in a fixed order at exit (before
# the HIP runtime is torn down); after shutdown() the finalizers do nothing.
_live = weakref.WeakSet()
_closed = False


def shutdown() -> None:
    """Frees every live object (batches, libraries, histograms, methods,
    pairs, tracks, device arrays, in that order), then calls st_shutdown,
    which waits for the library's streams and destroys them.  Registered to
    run at interpreter exit; objects must not be used afterwards."""
    global _closed
    if _closed:
        return
    objs = list(_live)
    for cls in (BatchRun, TrackLibrary, Histogram, VoxelConstrainedSphereMethod, VolumeEdepPair, Track,
                DeviceArray):
        for o in objs:
            if isinstance(o, cls):
                o.free()
    if _capi._lib is not None:
        check = _capi._lib.st_shutdown()
        if check != 0:
            raise SuperTrackError(check, "st_shutdown", _capi._lib.st_last_error().decode())
    _closed = True


atexit.register(shutdown)


def _finalize(obj) -> None:
    """__del__ body: frees the object unless the library has been shut down."""
    if _closed:
        return
    try:
        obj.free()
    except Exception:
        pass


def _dptr(a: np.ndarray, dtype) -> C.POINTER:
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(C.POINTER(np.ctypeslib.as_ctypes_type(dtype)))


def device_count() -> int:
    n = C.c_int32(0)
    call("st_device_count", C.byref(n))
    return int(n.value)


def synchronize() -> None:
    call("st_device_synchronize")


def default_stream_handle() -> C.c_void_p:
    """hipStream_t used by objects that were not given a stream."""
    s = C.c_void_p()
    call("st_default_stream", C.byref(s))
    return s


class SphericalGeometry:
    """SphericalGeometry (SphericalGeometry.cu:34-58)."""

    def __init__(self, half_length: float, diameter: float):
        self.c = _capi.Geometry()
        call("st_geometry_make", C.c_double(half_length), C.c_double(diameter), C.byref(self.c))

    half_length = property(lambda self: self.c.half_length)
    length = property(lambda self: self.c.length)
    diameter = property(lambda self: self.c.diameter)
    radius = property(lambda self: self.c.radius)
    gso = property(lambda self: self.c.gso)
    n_linear = property(lambda self: self.c.n_linear)

    def __repr__(self):
        return (f"SphericalGeometry(H'={self.half_length}, D={self.diameter}, N={self.n_linear}, "
                f"gso={self.gso})")


class DeviceArray:
    """A device allocation made through st_malloc."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        call("st_malloc", C.byref(self.ptr), C.c_size_t(max(self.nbytes, 1)))
        _live.add(self)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            call("st_memcpy_htod", d.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))
        return d

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            call("st_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().st_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        _finalize(self)


class Track:
    """Track: device SoA x, y, z (nm) and edep (eV) (Track.cuh:18-21)."""

    def __init__(self, x=None, y=None, z=None, edep=None, capacity: int = 0):
        self.c = _capi.Track()
        _live.add(self)
        if x is not None:
            self.upload(x, y, z, edep)
        else:
            call("st_track_alloc", C.byref(self.c), C.c_int64(capacity))

    def upload(self, x, y, z, edep):
        n = len(x)
        xs, px = _dptr(x, np.float64)
        ys, py = _dptr(y, np.float64)
        zs, pz = _dptr(z, np.float64)
        es, pe = _dptr(edep, np.float64)
        call("st_track_upload", C.byref(self.c), px, py, pz, pe, C.c_int64(n))

    def copy_from_device(self, src_ptrs: Sequence[int], n: int):
        """Device-to-device load of n points (the analogue of loading a track)."""
        if n > self.c.capacity:
            lib().st_track_free(C.byref(self.c))
            call("st_track_alloc", C.byref(self.c), C.c_int64(n))
        for dst, src in zip((self.c.x, self.c.y, self.c.z, self.c.edep), src_ptrs):
            call("st_memcpy_dtod", C.c_void_p(dst), C.c_void_p(src), C.c_size_t(8 * n))
        self.c.n = n

    @property
    def n(self) -> int:
        return int(self.c.n)

    def download(self):
        n = self.n
        out = []
        for p in (self.c.x, self.c.y, self.c.z, self.c.edep):
            a = np.empty(n, np.float64)
            if n:
                call("st_memcpy_dtoh", a.ctypes.data_as(C.c_void_p), C.c_void_p(p), C.c_size_t(8 * n))
            out.append(a)
        return tuple(out)

    def free(self):
        if self.c.x:
            lib().st_track_free(C.byref(self.c))

    def __del__(self):
        _finalize(self)


class VolumeEdepPair:
    """VolumeEdepPair: device (sphere key, lineal energy) pairs + count."""

    def __init__(self, capacity: int):
        self.c = _capi.Pairs()
        call("st_pairs_alloc", C.byref(self.c), C.c_int64(capacity))
        _live.add(self)

    @property
    def capacity(self) -> int:
        return int(self.c.capacity)

    def count(self) -> int:
        """*numElements, read on the host after the device has finished (as the
        reference reads its managed counter after the track's synchronisation)."""
        synchronize()
        return int(C.cast(self.c.num_elements, C.POINTER(C.c_int32))[0])

    def download(self):
        m = self.count()
        keys = np.empty(m, np.uint64)
        ys = np.empty(m, np.float64)
        if m:
            call("st_memcpy_dtoh", keys.ctypes.data_as(C.c_void_p), C.c_void_p(self.c.volume), C.c_size_t(8 * m))
            call("st_memcpy_dtoh", ys.ctypes.data_as(C.c_void_p), C.c_void_p(self.c.edep), C.c_size_t(8 * m))
        return keys, ys

    def upload(self, keys, ys):
        """Test helper: fill the pairs from host arrays (like HistogramTests' fill kernels)."""
        keys = np.ascontiguousarray(keys, np.uint64)
        ys = np.ascontiguousarray(ys, np.float64)
        m = len(keys)
        if m > self.capacity:
            raise ValueError("more pairs than capacity")
        if m:
            call("st_memcpy_htod", C.c_void_p(self.c.volume), keys.ctypes.data_as(C.c_void_p), C.c_size_t(8 * m))
            call("st_memcpy_htod", C.c_void_p(self.c.edep), ys.ctypes.data_as(C.c_void_p), C.c_size_t(8 * m))
        call("st_pairs_set_count", C.byref(self.c), C.c_int32(m))

    def free(self):
        if self.c.volume:
            lib().st_pairs_free(C.byref(self.c))

    def __del__(self):
        _finalize(self)


class VoxelConstrainedSphereMethod:
    """VoxelConstrainedSphereMethod (VoxelConstrainedSphereMethod.cu:9-119)."""

    def __init__(self, geometry: SphericalGeometry, shift_track: bool = True, shift_track_z: bool = False,
                 suggested_blocks: int = 0, suggested_threads: int = 0):
        self.geometry = geometry
        self.h = C.c_void_p()
        call("st_vcs_create", C.byref(geometry.c), int(bool(shift_track)), int(bool(shift_track_z)),
             int(suggested_blocks), int(suggested_threads), C.byref(self.h))
        _live.add(self)

    def allocate_track_process(self, track: Track, seed: int, track_index: int, n_oversamples: int):
        call("st_vcs_begin_track", self.h, C.byref(track.c), C.c_uint64(seed), C.c_uint32(track_index),
             C.c_int32(n_oversamples))

    def process_track(self, track: Track, pairs: VolumeEdepPair):
        call("st_vcs_process_oversample", self.h, C.byref(track.c), C.byref(pairs.c))

    def free_track_process(self):
        call("st_vcs_end_track", self.h)

    def set_stream(self, stream: C.c_void_p):
        """Runs this method's work on a given hipStream_t (one per worker)."""
        call("st_vcs_set_stream", self.h, stream)

    def set_queued(self, on: bool = True):
        """Queued oversamples: a track's oversamples run as one launch when the
        track ends (st_vcs_set_queued); the pairs are defined only after that."""
        call("st_vcs_set_queued", self.h, int(bool(on)))

    def free(self):
        if self.h:
            lib().st_vcs_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        _finalize(self)


class Histogram:
    """Histogram (Histogram.cu:16-199) with a ROOT-free CPU copy.

    cpu holds nbins + 2 doubles in TH1D order: 0 underflow, 1..nbins, nbins + 1 overflow.
    """

    def __init__(self, nbins: int, lo: float, hi: float, kind: str = "lin"):
        kinds = {"lin": ST_HIST_LIN, "log": ST_HIST_LOG}
        if kind not in kinds:
            raise ValueError(f"Unknown histogram type: {kind} available types are lin and log.")
        self.kind = kind
        self.h = C.c_void_p()
        call("st_hist_create", kinds[kind], C.c_float(lo), C.c_float(hi), C.c_int32(nbins), C.byref(self.h))
        _live.add(self)
        self.nbins = int(nbins)
        self.edges = np.empty(self.nbins + 1, np.float64)
        call("st_hist_edges", self.h, self.edges.ctypes.data_as(_capi.c_double_p))
        self.cpu = np.zeros(self.nbins + 2, np.float64)

    def allocate_track_process(self, pairs: VolumeEdepPair):
        call("st_hist_begin_track", self.h, C.byref(pairs.c))

    def sort_reduce_and_add_to_histogram(self, pairs: VolumeEdepPair):
        call("st_hist_add", self.h, C.byref(pairs.c))

    def free_track_process(self):
        call("st_hist_end_track", self.h, self.cpu.ctypes.data_as(_capi.c_double_p))

    def set_stream(self, stream: C.c_void_p):
        """Runs this histogram's work on a given hipStream_t (one per worker)."""
        call("st_hist_set_stream", self.h, stream)

    def fetch(self, reset: bool = True) -> np.ndarray:
        out = np.zeros(self.nbins, np.uint64)
        call("st_hist_fetch", self.h, out.ctypes.data_as(_capi.c_uint64_p), int(bool(reset)))
        return out

    def counts(self) -> np.ndarray:
        """Bins 1..nbins of the CPU copy."""
        return self.cpu[1:self.nbins + 1].copy()

    def free(self):
        if self.h:
            lib().st_hist_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        _finalize(self)


class TrackLibrary:
    """A device-resident track library laid out for the batched kernel."""

    def __init__(self, x, y, z, e, offsets, geometries: Sequence[SphericalGeometry],
                 geom_of_track: Optional[np.ndarray] = None, precision: int = ST_F64):
        xs, px = _dptr(x, np.float64)
        ys, py = _dptr(y, np.float64)
        zs, pz = _dptr(z, np.float64)
        es, pe = _dptr(e, np.float64)
        offs, po = _dptr(offsets, np.int64)
        n_tracks = len(offs) - 1
        garr = (_capi.Geometry * len(geometries))(*[g.c for g in geometries])
        if geom_of_track is not None:
            gt, pg = _dptr(geom_of_track, np.int32)
        else:
            gt, pg = None, None
        self.h = C.c_void_p()
        call("st_library_create", px, py, pz, pe, po, C.c_int64(n_tracks), garr, len(geometries), pg,
             int(precision), C.byref(self.h))
        _live.add(self)
        self.geometries = list(geometries)

    @classmethod
    def streamed(cls, n_tracks: int, n_points: int, chunks: Iterable, geometries: Sequence[SphericalGeometry],
                 precision: int = ST_F64) -> "TrackLibrary":
        """Builds the library from an iterable of (x, y, z, e, offsets, geom_of_track)
        chunks in library order, holding one chunk on the host at a time;
        n_tracks and n_points are the totals over all chunks."""
        self = cls.__new__(cls)
        self.h = C.c_void_p()
        self.geometries = list(geometries)
        garr = (_capi.Geometry * len(geometries))(*[g.c for g in geometries])
        call("st_library_begin", C.c_int64(n_tracks), C.c_int64(n_points), garr, len(geometries), int(precision),
             C.byref(self.h))
        _live.add(self)
        for x, y, z, e, offsets, gt in chunks:
            xs, px = _dptr(x, np.float64)
            ys, py = _dptr(y, np.float64)
            zs, pz = _dptr(z, np.float64)
            es, pe = _dptr(e, np.float64)
            offs, po = _dptr(offsets, np.int64)
            if gt is not None:
                gts, pg = _dptr(gt, np.int32)
            else:
                gts, pg = None, None
            call("st_library_append", self.h, px, py, pz, pe, po, C.c_int64(len(offs) - 1), pg)
        call("st_library_finish", self.h)
        return self

    def info(self) -> dict:
        i = _capi.LibraryInfo()
        call("st_library_info_get", self.h, C.byref(i))
        return {f: getattr(i, f) for f, _ in i._fields_}

    def free(self):
        if self.h:
            lib().st_library_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        _finalize(self)


class BatchRun:
    """One fused persistent pass over every track x oversample of a library."""

    def __init__(self, library: TrackLibrary, histograms: Sequence[Histogram], n_oversamples: int,
                 seed: int, shift_xy: bool = True, shift_z: bool = False, track_index_base: int = 0,
                 use_graph: bool = True, autotune: bool = False, block_threads: int = 0,
                 table_slots: int = 0, blocks_per_cu: int = 0, points_per_thread: int = 0,
                 hist_copies: int = 0):
        o = _capi.BatchOpts()
        o.n_oversamples = n_oversamples
        o.shift_xy = int(bool(shift_xy))
        o.shift_z = int(bool(shift_z))
        o.track_index_base = track_index_base
        o.seed = seed
        o.use_graph = int(bool(use_graph))
        o.autotune = int(bool(autotune))
        o.block_threads = block_threads
        o.table_slots = table_slots
        o.blocks_per_cu = blocks_per_cu
        o.points_per_thread = points_per_thread
        o.hist_copies = hist_copies
        hs = (C.c_void_p * len(histograms))(*[h.h for h in histograms])
        self.h = C.c_void_p()
        self.library = library
        self.histograms = list(histograms)
        call("st_batch_create", library.h, C.byref(o), hs, len(histograms), C.byref(self.h))
        _live.add(self)

    def run(self):
        call("st_batch_run", self.h)

    def sync(self):
        call("st_batch_sync", self.h)

    def times(self):
        """(fused kernel ms, exact path ms) of the last run."""
        t = (C.c_float * 2)()
        call("st_batch_times", self.h, t)
        return float(t[0]), float(t[1])

    def stats(self):
        """Point counts of the last run: scored, FP64 rechecks, inside a sphere."""
        c = np.zeros(3, np.int64)
        call("st_batch_stats", self.h, c.ctypes.data_as(_capi.c_int64_p))
        return {"points_scored": int(c[0]), "fp64_rechecks": int(c[1]), "points_in_sphere": int(c[2])}

    def balance(self):
        """Conservation totals of the last run (modulo 2^64): energy words into
        and out of the sphere tables, entries created and retired."""
        c = np.zeros(4, np.uint64)
        call("st_batch_balance", self.h, c.ctypes.data_as(_capi.c_uint64_p))
        return {"energy_in": int(c[0]), "energy_out": int(c[1]), "entries_created": int(c[2]),
                "entries_retired": int(c[3])}

    def params(self):
        p = np.zeros(5, np.int32)
        call("st_batch_params", self.h, p.ctypes.data_as(_capi.c_int32_p))
        return {"block_threads": int(p[0]), "table_slots": int(p[1]), "blocks_per_cu": int(p[2]),
                "points_per_thread": int(p[3]), "hist_copies": int(p[4])}

    def record(self, capacity: int):
        call("st_batch_record", self.h, C.c_int64(capacity))

    def record_fetch(self, capacity: int):
        seg = np.zeros(capacity, np.int64)
        keys = np.zeros(capacity, np.uint64)
        sums = np.zeros(capacity, np.float64)
        cnt = C.c_int64(0)
        call("st_batch_record_fetch", self.h, seg.ctypes.data_as(_capi.c_int64_p),
             keys.ctypes.data_as(_capi.c_uint64_p), sums.ctypes.data_as(_capi.c_double_p), C.byref(cnt))
        n = int(cnt.value)
        return seg[:n], keys[:n], sums[:n]

    def free(self):
        if self.h:
            lib().st_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        _finalize(self)


def synth_counts(spec: dict, n_tracks: int, t0: int = 0) -> np.ndarray:
    """Point counts of tracks [t0, t0 + n_tracks) of a synthetic spec, without generating them."""
    s = _capi.SynthSpec()
    for k, v in spec.items():
        setattr(s, k, v)
    counts = np.zeros(n_tracks, np.int64)
    call("st_synth_counts", C.byref(s), C.c_int64(t0), C.c_int64(t0 + n_tracks),
         counts.ctypes.data_as(_capi.c_int64_p))
    return counts


def synth_library(spec: dict, n_tracks: int, t0: int = 0):
    """Generates tracks [t0, t0 + n_tracks) of a synthetic spec.

    Returns (x, y, z, edep, offsets) as numpy arrays (offsets has n_tracks + 1 entries).
    """
    s = _capi.SynthSpec()
    for k, v in spec.items():
        setattr(s, k, v)
    counts = synth_counts(spec, n_tracks, t0)
    offsets = np.zeros(n_tracks + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    n = int(offsets[-1])
    x = np.empty(n, np.float64)
    y = np.empty(n, np.float64)
    z = np.empty(n, np.float64)
    e = np.empty(n, np.float64)
    call("st_synth_fill", C.byref(s), C.c_int64(t0), C.c_int64(t0 + n_tracks),
         offsets.ctypes.data_as(_capi.c_int64_p), x.ctypes.data_as(_capi.c_double_p),
         y.ctypes.data_as(_capi.c_double_p), z.ctypes.data_as(_capi.c_double_p),
         e.ctypes.data_as(_capi.c_double_p))
    return x, y, z, e, offsets


def run_plain(x, y, z, e, offsets, geometry: SphericalGeometry, histogram: Histogram, n_oversamples: int,
              seed: int, shift_xy: bool = True, shift_z: bool = False, track_index_base: int = 0,
              tracks: Optional[Iterable[int]] = None, queued: bool = False):
    """Per-track loop of SuperTrackManager.cu:58-83 over host arrays (the plain drop-in path).

    Counts are added to histogram.cpu.  queued: the method queues each track's
    oversamples and runs them in one launch at the track's end.  Returns the
    number of tracks processed.
    """
    method = VoxelConstrainedSphereMethod(geometry, shift_xy, shift_z)
    if queued:
        method.set_queued(True)
    n_tracks = len(offsets) - 1
    done = 0
    for t in (range(n_tracks) if tracks is None else tracks):
        b, en = int(offsets[t]), int(offsets[t + 1])
        track = Track(x[b:en], y[b:en], z[b:en], e[b:en])
        pairs = VolumeEdepPair(en - b)
        method.allocate_track_process(track, seed, track_index_base + t, n_oversamples)
        histogram.allocate_track_process(pairs)
        for _ in range(n_oversamples):
            method.process_track(track, pairs)
            histogram.sort_reduce_and_add_to_histogram(pairs)
        method.free_track_process()
        histogram.free_track_process()
        track.free()
        pairs.free()
        done += 1
    method.free()
    return done
