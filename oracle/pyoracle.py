"""ctypes view of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time the FP64 restatement of the
reference; the product (supertrack_amd) never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SUPERTRACK_ORACLE_LIBRARY") or os.path.join(HERE, "liboracle.so")
PHILOX_CHECK = os.path.join(HERE, "philox_rocrand_check")

_d = C.POINTER(C.c_double)
_u64 = C.POINTER(C.c_uint64)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)
_u32 = C.POINTER(C.c_uint32)


class OrcGeometry(C.Structure):
    _fields_ = [("half_length", C.c_double), ("length", C.c_double), ("diameter", C.c_double),
                ("radius", C.c_double), ("gso", C.c_double), ("n_linear", C.c_int32), ("pad", C.c_int32)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    L.orc_geometry_make.argtypes = [C.c_double, C.c_double, C.POINTER(OrcGeometry)]
    L.orc_geometry_make.restype = None
    L.orc_philox4x32_10.argtypes = [_u32, _u32, _u32]
    L.orc_philox4x32_10.restype = None
    L.orc_u01_open_closed.argtypes = [C.c_uint32]
    L.orc_u01_open_closed.restype = C.c_float
    L.orc_shift_table.argtypes = [C.c_uint64, C.c_uint32, C.c_int32, C.c_int32, C.c_int32,
                                  C.POINTER(OrcGeometry), _d]
    L.orc_shift_table.restype = None
    L.orc_edges.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_int32, _d]
    L.orc_edges.restype = None
    L.orc_bin.argtypes = [_d, C.c_int32, C.c_double]
    L.orc_bin.restype = C.c_int32
    L.orc_track_zshift.argtypes = [_d, C.c_int64, C.c_double]
    L.orc_track_zshift.restype = None
    L.orc_score_oversample.argtypes = [_d, _d, _d, _d, C.c_int64, C.POINTER(OrcGeometry), C.c_double,
                                       C.c_double, C.c_double, _u64, _d, _i64, _u64, _d]
    L.orc_score_oversample.restype = C.c_int64
    L.orc_reduce_pairs.argtypes = [_u64, _d, C.c_int64, _u64, _d]
    L.orc_reduce_pairs.restype = C.c_int64
    L.orc_run.argtypes = [_d, _d, _d, _d, _i64, C.c_int64, C.c_int64, C.POINTER(OrcGeometry), _i32,
                          C.c_int32, C.c_uint64, C.c_uint32, C.c_int32, C.c_int32, _d, C.c_int32, _u64,
                          _d, C.c_int32, _u64, C.c_int32, _i32, C.c_double, _u64]
    L.orc_near_edge.argtypes = [_d, C.c_int32, C.c_double, C.c_double]
    L.orc_near_edge.restype = C.c_int32
    L.orc_run.restype = C.c_double
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def geometry(half_length: float, diameter: float) -> OrcGeometry:
    g = OrcGeometry()
    lib().orc_geometry_make(half_length, diameter, C.byref(g))
    return g


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c, _u32), _p(k, _u32), _p(out, _u32))
    return out


def u01(r: int) -> float:
    return float(lib().orc_u01_open_closed(r))


def shift_table(seed: int, track_index: int, n_os: int, shift_xy: bool, shift_z: bool, g: OrcGeometry):
    out = np.zeros(3 * n_os, np.float64)
    lib().orc_shift_table(seed, track_index, n_os, int(shift_xy), int(shift_z), C.byref(g), _p(out, _d))
    return out.reshape(n_os, 3)


def edges(kind: str, lo: float, hi: float, nbins: int) -> np.ndarray:
    out = np.zeros(nbins + 1, np.float64)
    lib().orc_edges(1 if kind == "log" else 0, lo, hi, nbins, _p(out, _d))
    return out


def bin_index(e: np.ndarray, v: float) -> int:
    e = np.ascontiguousarray(e, np.float64)
    return int(lib().orc_bin(_p(e, _d), len(e) - 1, v))


def score_oversample(x, y, zp, e, g: OrcGeometry, sx: float, sy: float, sz: float):
    """(pair_keys, pair_y, keys, sums) of one oversample; zp is z + H'."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    zp = np.ascontiguousarray(zp, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    n = len(x)
    pk = np.zeros(max(n, 1), np.uint64)
    py = np.zeros(max(n, 1), np.float64)
    keys = np.zeros(max(n, 1), np.uint64)
    sums = np.zeros(max(n, 1), np.float64)
    m = C.c_int64(0)
    u = lib().orc_score_oversample(_p(x, _d), _p(y, _d), _p(zp, _d), _p(e, _d), n, C.byref(g), sx, sy, sz,
                                   _p(pk, _u64), _p(py, _d), C.byref(m), _p(keys, _u64), _p(sums, _d))
    return pk[:m.value], py[:m.value], keys[:u], sums[:u]


def reduce_pairs(pair_keys, pair_y):
    """(keys, sums) of Histogram's reduce-by-key in canonical order."""
    pk = np.ascontiguousarray(pair_keys, np.uint64)
    py = np.ascontiguousarray(pair_y, np.float64)
    m = len(pk)
    keys = np.zeros(max(m, 1), np.uint64)
    sums = np.zeros(max(m, 1), np.float64)
    u = lib().orc_reduce_pairs(_p(pk, _u64), _p(py, _d), m, _p(keys, _u64), _p(sums, _d))
    return keys[:u], sums[:u]


def histogram_counts(e: np.ndarray, values) -> np.ndarray:
    """Counts of values in the bins of edges e (HistogramRange semantics)."""
    out = np.zeros(len(e) - 1, np.uint64)
    for v in values:
        b = bin_index(e, float(v))
        if b >= 0:
            out[b] += 1
    return out


def track_segments(x, y, z, e, g: OrcGeometry, seed: int, track_index: int, n_os: int, shift_xy=True,
                   shift_z=False):
    """List over oversamples of (pair_keys, pair_y, keys, sums) for one track in its own order."""
    zp = np.ascontiguousarray(z, np.float64).copy()
    lib().orc_track_zshift(_p(zp, _d), len(zp), g.half_length)
    sh = shift_table(seed, track_index, n_os, shift_xy, shift_z, g)
    return [score_oversample(x, y, zp, e, g, *sh[k]) for k in range(n_os)]


def near_edge(e: np.ndarray, v: float, rtol: float) -> bool:
    """True if an edge adjacent to v lies within rtol * |v| of it."""
    e = np.ascontiguousarray(e, np.float64)
    return bool(lib().orc_near_edge(_p(e, _d), len(e) - 1, v, rtol))


def run(x, y, z, e, offsets, geoms, geom_of_track, n_os, seed, shift_xy=True, shift_z=False,
        hists=(), track_stride=1, track_index_base=0, n_threads=0, near_rtol=None):
    """Whole-run driver: returns ([bins per hist], seconds, threads), and with
    near_rtol set also [near-edge sums per hist] (orc_near_edge)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    offsets = np.ascontiguousarray(offsets, np.int64)
    garr = (OrcGeometry * len(geoms))(*geoms)
    gt = None if geom_of_track is None else np.ascontiguousarray(geom_of_track, np.int32)
    hs = list(hists) + [None] * (2 - len(hists))
    bins = [None if h is None else np.zeros(len(h) - 1, np.uint64) for h in hs]
    ed = [None if h is None else np.ascontiguousarray(h, np.float64) for h in hs]
    used = C.c_int32(0)
    near = np.zeros(2, np.uint64)
    secs = lib().orc_run(
        _p(x, _d), _p(y, _d), _p(z, _d), _p(e, _d), _p(offsets, _i64), len(offsets) - 1, track_stride, garr,
        None if gt is None else _p(gt, _i32), n_os, seed, track_index_base, int(shift_xy), int(shift_z),
        None if ed[0] is None else _p(ed[0], _d), 0 if ed[0] is None else len(ed[0]) - 1,
        None if bins[0] is None else _p(bins[0], _u64),
        None if ed[1] is None else _p(ed[1], _d), 0 if ed[1] is None else len(ed[1]) - 1,
        None if bins[1] is None else _p(bins[1], _u64), n_threads, C.byref(used),
        0.0 if near_rtol is None else float(near_rtol), None if near_rtol is None else _p(near, _u64))
    out = [b for b in bins if b is not None]
    if near_rtol is None:
        return out, float(secs), int(used.value)
    return out, float(secs), int(used.value), [int(near[h]) for h in range(len(out))]


def rocrand_philox(rows):
    """Runs the rocRAND-header Philox on rows of (c0, c1, c2, c3, k0, k1)."""
    if not os.path.exists(PHILOX_CHECK):
        build()
    text = "\n".join(" ".join(f"{v:08x}" for v in r) for r in rows) + "\n"
    out = subprocess.run([PHILOX_CHECK], input=text, capture_output=True, text=True, check=True).stdout
    return [tuple(int(w, 16) for w in line.split()) for line in out.strip().splitlines()]
