"""The five benchmark workloads of the VoxelConstrainedSphere path.

Each workload names its synthetic track library, geometry(ies), histogram(s),
oversample count and shift settings.  Sizes are the full ones; tests pass
smaller track counts through ``make_library(..., n_tracks=...)``.

  cfg1  1 track of ~4.9e3 points, FP64, 10 oversamples, lin 3000 bins
  cfg2  256 tracks of 2e3-1e4 points, 100 oversamples, plain FP64 path
  cfg3  4096 tracks, log-uniform 1e3-1e6 points, 100 oversamples, log 1000 bins, batched FP64
  cfg4  cfg3's library in FP32, D = 2 nm (N = 2500: reference keys beyond 32 bits), 1000 oversamples,
        batched FP32
  cfg5  16384 tracks, 1e3-2e6 points, three geometries by 1024-track chunk, 1000 oversamples,
        lin 3000 + log 1000 bins
"""
from __future__ import annotations

import copy

import numpy as np

from . import ST_F32, ST_F64, SphericalGeometry, synth_counts, synth_library

SEED = 1234

_BASE_SYNTH = dict(
    voxel_half_length=2500.0, entry_half_length=2500.0,
    length_min_nm=10000.0, length_max_nm=10000.0, spacing_nm=14.0,
    p_cluster=0.3, p_cluster_max=0.0, cluster_mean=20.0, walk_sigma_nm=2.0,
    edep_median_ev=15.0, edep_sigma_ln=0.8, edep_min_ev=1.0, edep_max_ev=1000.0,
    spacing_min_nm=0.5, spacing_max_nm=50.0, n_min=0, n_max=0, stream_id=1,
)

LIN3000 = ("lin", 3000, 0.01, 300.0)
LOG1000 = ("log", 1000, 1e-2, 1e3)

WORKLOADS = {
    "cfg1": dict(
        n_tracks=1, geometries=[(2.5e3, 10.0)], chunk=0, hists=[LIN3000], n_oversamples=10,
        synth=dict(_BASE_SYNTH, stream_id=1), precision=ST_F64, path="plain",
    ),
    "cfg2": dict(
        n_tracks=256, geometries=[(2.5e3, 10.0)], chunk=0, hists=[LIN3000], n_oversamples=100,
        synth=dict(_BASE_SYNTH, length_min_nm=5000.0, length_max_nm=20000.0, stream_id=2), precision=ST_F64,
        path="plain",
    ),
    "cfg3": dict(
        n_tracks=4096, geometries=[(2.5e3, 10.0)], chunk=0, hists=[LOG1000], n_oversamples=100,
        synth=dict(_BASE_SYNTH, length_min_nm=5000.0, length_max_nm=20000.0, p_cluster=0.1,
                   p_cluster_max=0.6, n_min=1000, n_max=1000000, stream_id=3),
        precision=ST_F64,
    ),
    "cfg4": dict(
        n_tracks=4096, geometries=[(2.5e3, 2.0)], chunk=0, hists=[LOG1000], n_oversamples=1000,
        synth=dict(_BASE_SYNTH, length_min_nm=5000.0, length_max_nm=20000.0, p_cluster=0.1,
                   p_cluster_max=0.6, n_min=1000, n_max=1000000, stream_id=3),
        precision=ST_F32,
    ),
    "cfg5": dict(
        n_tracks=16384, geometries=[(1e3, 1.0), (2.5e3, 10.0), (5e3, 100.0)], chunk=1024,
        hists=[LIN3000, LOG1000], n_oversamples=1000,
        synth=dict(_BASE_SYNTH, length_min_nm=5000.0, length_max_nm=20000.0, p_cluster=0.1,
                   p_cluster_max=0.6, n_min=1000, n_max=2000000, stream_id=5),
        precision=ST_F32,
    ),
}


def workload(name: str) -> dict:
    return copy.deepcopy(WORKLOADS[name])


def geometries(w: dict):
    return [SphericalGeometry(h, d) for h, d in w["geometries"]]


def geom_of_track(w: dict, n_tracks: int, t0: int = 0) -> np.ndarray:
    """Geometry index per track: chunks of w['chunk'] tracks cycle through the geometries."""
    ng = len(w["geometries"])
    if ng == 1 or not w["chunk"]:
        return np.zeros(n_tracks, np.int32)
    t = np.arange(t0, t0 + n_tracks, dtype=np.int64)
    return ((t // w["chunk"]) % ng).astype(np.int32)


def _geometry_runs(w: dict, n_tracks: int, max_tracks: int):
    """(first track, count, geometry) of consecutive same-geometry runs of at most max_tracks tracks."""
    g_of_t = geom_of_track(w, n_tracks)
    runs = []
    t = 0
    while t < n_tracks:
        u = t + 1
        while u < n_tracks and u - t < max_tracks and g_of_t[u] == g_of_t[t]:
            u += 1
        runs.append((t, u - t, int(g_of_t[t])))
        t = u
    return runs


def _spec_for(w: dict, g: SphericalGeometry) -> dict:
    spec = dict(w["synth"])
    spec["entry_half_length"] = g.half_length
    spec["voxel_half_length"] = g.half_length
    return spec


def library_chunks(w: dict, n_tracks: int = None, max_tracks: int = 1024, only_geometry: int = None):
    """The library of make_library(w, n_tracks) as (n_tracks, n_points, chunks): chunks
    yields (x, y, z, edep, offsets, geom_of_track) per run of at most max_tracks
    consecutive same-geometry tracks, generated when it is consumed, so only one
    run is held on the host at a time.  With only_geometry = g, only the tracks
    of geometry g, as a one-geometry library (geometry index 0)."""
    n_tracks = w["n_tracks"] if n_tracks is None else n_tracks
    geoms = geometries(w)
    runs = _geometry_runs(w, n_tracks, max_tracks)
    if only_geometry is not None:
        runs = [r for r in runs if r[2] == only_geometry]
    n_points = int(sum(int(synth_counts(_spec_for(w, geoms[g]), c, t0).sum()) for t0, c, g in runs))

    def gen():
        for t0, c, g in runs:
            x, y, z, e, off = synth_library(_spec_for(w, geoms[g]), c, t0)
            yield x, y, z, e, off, np.full(c, 0 if only_geometry is not None else g, np.int32)

    return sum(c for _, c, _ in runs), n_points, gen()


def make_library(w: dict, n_tracks: int = None, t0: int = 0, stride: int = 1):
    """Generates tracks t0, t0 + stride, ... of a workload's library.

    Returns (x, y, z, edep, offsets, geom_of_track).  The primary of each track
    enters the voxel of its own geometry (entry and x/y range follow H').
    """
    n_tracks = w["n_tracks"] if n_tracks is None else n_tracks
    geoms = geometries(w)
    g_of_t = geom_of_track(w, n_tracks * stride, t0)[::stride]
    idx = t0 + stride * np.arange(n_tracks)
    parts = []
    if stride == 1 and len(geoms) == 1:
        spec = dict(w["synth"])
        spec["entry_half_length"] = geoms[0].half_length
        spec["voxel_half_length"] = geoms[0].half_length
        x, y, z, e, off = synth_library(spec, n_tracks, t0)
        return x, y, z, e, off, g_of_t
    # per track (strided subsets or mixed geometries): generate one at a time, grouped by runs
    xs, ys, zs, es, counts = [], [], [], [], []
    for j, t in enumerate(idx):
        g = geoms[int(g_of_t[j])]
        spec = dict(w["synth"])
        spec["entry_half_length"] = g.half_length
        spec["voxel_half_length"] = g.half_length
        x, y, z, e, off = synth_library(spec, 1, int(t))
        xs.append(x); ys.append(y); zs.append(z); es.append(e); counts.append(len(x))
    off = np.zeros(n_tracks + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    return (np.concatenate(xs) if xs else np.zeros(0), np.concatenate(ys) if ys else np.zeros(0),
            np.concatenate(zs) if zs else np.zeros(0), np.concatenate(es) if es else np.zeros(0), off, g_of_t)
