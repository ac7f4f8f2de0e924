"""Parity of the batched persistent kernel with the FP64 oracle.

Sphere membership (the set of keys of every (track, oversample) segment) must
match exactly; per-sphere energies within 1e-12 relative (ST_F64: exact
fixed-point sums vs the oracle's left-to-right sums) or 1e-6 (ST_F32); every
histogram count must match except for spheres whose oracle energy lies
within that tolerance of a bin edge (counted and reported)."""
import numpy as np
import pytest

import fixtures as F
import pyoracle as po
import supertrack_amd as st
from supertrack_amd import workloads as W

pytestmark = pytest.mark.gpu


def oracle_segments(x, y, z, e, off, geoms, gt, n_os, seed, shift_xy, shift_z):
    out = {}
    for t in range(len(off) - 1):
        b, en = off[t], off[t + 1]
        g = geoms[int(gt[t]) if gt is not None else 0]
        segs = po.track_segments(x[b:en], y[b:en], z[b:en], e[b:en], g, seed, t, n_os, shift_xy, shift_z)
        for os, (_, _, keys, sums) in enumerate(segs):
            out[t * n_os + os] = (keys, sums)
    return out


def gpu_run(x, y, z, e, off, geoms_hd, gt, n_os, seed, shift_xy, shift_z, precision, hists, record=0, **kw):
    geoms = [st.SphericalGeometry(*g) for g in geoms_hd]
    lib = st.TrackLibrary(x, y, z, e, off, geoms, gt, precision)
    hs = [st.Histogram(nb, lo, hi, kind) for kind, nb, lo, hi in hists]
    b = st.BatchRun(lib, hs, n_os, seed, shift_xy, shift_z, **kw)
    if record:
        b.record(record)
    b.run()
    b.sync()
    rec = b.record_fetch(record) if record else None
    counts = [h.fetch() for h in hs]
    info = lib.info()
    info.update(b.stats())
    bal = b.balance()
    # conservation: every energy word and entry that entered a sphere table left it
    assert bal["energy_in"] == bal["energy_out"] and bal["entries_created"] == bal["entries_retired"], bal
    assert (bal["entries_created"] > 0) == (info["points_in_sphere"] > 0)
    edges = [h.edges.copy() for h in hs]
    b.free()
    for h in hs:
        h.free()
    lib.free()
    return counts, rec, info, edges


def compare_records(rec, oracle, rtol):
    """Keys per segment identical; sums within rtol.  Returns (n spheres, max rel err)."""
    seg, keys, sums = rec
    order = np.lexsort((keys, seg))
    seg, keys, sums = seg[order], keys[order], sums[order]
    got = {}
    if len(seg):
        cuts = np.flatnonzero(np.diff(seg)) + 1
        for s_idx in np.split(np.arange(len(seg)), cuts):
            got[int(seg[s_idx[0]])] = (keys[s_idx], sums[s_idx])
    n, worst = 0, 0.0
    for s, (wk, ws) in oracle.items():
        gk, gs = got.get(s, (np.zeros(0, np.uint64), np.zeros(0)))
        np.testing.assert_array_equal(gk, wk, err_msg=f"segment {s}")
        if len(ws):
            rel = np.abs(gs - ws) / np.maximum(np.abs(ws), 1e-300)
            worst = max(worst, float(rel.max()))
            assert rel.max() <= rtol, (s, rel.max())
        n += len(wk)
    assert set(got) <= set(oracle)
    return n, worst


def check_bins(counts, edges, oracle, rtol):
    """Histogram counts equal the oracle's except for near-edge spheres.
    Returns (count differences, near-edge sums) summed over the histograms."""
    sums = np.concatenate([s for _, s in oracle.values()]) if oracle else np.zeros(0)
    tot_diff, tot_near = 0, 0
    for c, ed in zip(counts, edges):
        want = po.histogram_counts(ed, sums)
        diff = np.abs(c.astype(np.int64) - want.astype(np.int64)).sum()
        # spheres whose oracle energy is within rtol of an edge may flip bins
        idx = np.searchsorted(ed, sums)
        near = 0
        for v, i in zip(sums, idx):
            for j in (i - 1, i):
                if 0 <= j < len(ed) and abs(v - ed[j]) <= rtol * abs(v):
                    near += 1
                    break
        assert diff <= 2 * near, (diff, near)
        tot_diff += int(diff)
        tot_near += near
    return tot_diff, tot_near


@pytest.mark.parametrize("precision,rtol", [(st.ST_F64, 1e-12), (st.ST_F32, 1e-6)])
def test_batch_cfg3_like_subset(precision, rtol):
    w = W.workload("cfg3")
    spec = dict(w["synth"], n_max=60000)   # keep the oracle fast
    x, y, z, e, off = st.synth_library(spec, 16, 0)
    n_os = 6
    geoms = [po.geometry(2.5e3, 10.0)]
    oracle = oracle_segments(x, y, z, e, off, geoms, None, n_os, W.SEED, True, True)
    counts, rec, info, edges = gpu_run(x, y, z, e, off, [(2.5e3, 10.0)], None, n_os, W.SEED, True, True,
                                       precision, [("log", 1000, 1e-2, 1e3)], record=400000,
                                       points_per_thread=1 + precision)
    assert info["exact_tracks"] == 0
    n, worst = compare_records(rec, oracle, rtol)
    assert n > 10000
    check_bins(counts, edges, oracle, rtol)


@pytest.mark.parametrize("name,precision,rtol,stride,n_os", [
    ("cfg3", st.ST_F64, 1e-12, 256, 20),
    ("cfg3", st.ST_F32, 1e-6, 256, 20),
    ("cfg4", st.ST_F32, 1e-6, 256, 4),
    ("cfg5", st.ST_F32, 1e-6, 1024, 10),
])
def test_batch_parity_report_per_config(name, precision, rtol, stride, n_os):
    # every stride-th track of the workload's library (full sizes, up to 1e6 or
    # 2e6 points), all its geometries and histograms: sphere keys exact, the
    # worst per-sphere energy error, and how many sums lie within the tolerance
    # of a bin edge (the only counts allowed to differ)
    w = W.workload(name)
    n_sample = w["n_tracks"] // stride
    x, y, z, e, off, gt = W.make_library(w, n_tracks=n_sample, stride=stride)
    geo = w["geometries"]
    mixed = len(geo) > 1
    geoms = [po.geometry(*g) for g in geo]
    oracle = oracle_segments(x, y, z, e, off, geoms, gt if mixed else None, n_os, W.SEED, True, True)
    n_rec = sum(len(k) for k, _ in oracle.values())
    counts, rec, info, edges = gpu_run(x, y, z, e, off, geo, gt if mixed else None, n_os, W.SEED, True, True,
                                       precision, w["hists"], record=n_rec + 1024)
    assert info["exact_tracks"] == 0
    n, worst = compare_records(rec, oracle, rtol)
    diff, near = check_bins(counts, edges, oracle, rtol)
    print(f"\nPARITY {name} {'F64' if precision == st.ST_F64 else 'F32'}: {n_sample} tracks "
          f"({int(off[-1])} points) x {n_os} oversamples, {n} spheres, max relative energy error "
          f"{worst:.3e} (tolerance {rtol:g}), {near} sums within tolerance of a bin edge, "
          f"{diff} bin counts differ", flush=True)
    assert n > 100000


def test_batch_f32_small_spheres_64bit_keys():
    # cfg4 geometry: D = 2 nm, N = 2500, keys above 2^32
    w = W.workload("cfg4")
    spec = dict(w["synth"], n_max=40000)
    x, y, z, e, off = st.synth_library(spec, 10, 100)
    geoms = [po.geometry(2.5e3, 2.0)]
    oracle = oracle_segments(x, y, z, e, off, geoms, None, 5, 77, True, True)
    assert max(int(k.max()) for k, _ in oracle.values() if len(k)) > 2**32
    counts, rec, info, edges = gpu_run(x, y, z, e, off, [(2.5e3, 2.0)], None, 5, 77, True, True, st.ST_F32,
                                       [("log", 1000, 1e-2, 1e3)], record=400000)
    compare_records(rec, oracle, 1e-6)
    check_bins(counts, edges, oracle, 1e-6)


def test_batch_mixed_geometries_two_histograms():
    w = W.workload("cfg5")
    w["chunk"] = 2
    spec = dict(w["synth"], n_max=30000)
    geo = w["geometries"]
    xs, ys, zs, es, cnt = [], [], [], [], []
    gt = W.geom_of_track(w, 12)
    for t in range(12):
        H, D = geo[gt[t]]
        g = st.SphericalGeometry(H, D)
        x, y, z, e, o = st.synth_library(dict(spec, entry_half_length=g.half_length,
                                              voxel_half_length=g.half_length), 1, t)
        xs.append(x); ys.append(y); zs.append(z); es.append(e); cnt.append(len(x))
    off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    x, y, z, e = (np.concatenate(a) for a in (xs, ys, zs, es))
    geoms = [po.geometry(*g) for g in geo]
    oracle = oracle_segments(x, y, z, e, off, geoms, gt, 4, 5, True, True)
    counts, rec, info, edges = gpu_run(x, y, z, e, off, geo, gt, 4, 5, True, True, st.ST_F32,
                                       [("lin", 3000, 0.01, 300.0), ("log", 1000, 1e-2, 1e3)], record=400000)
    # tracks routed to the exact path are not recorded: compare recorded ones only
    exact = set()
    if info["exact_tracks"]:
        seen = {int(s) // 4 for s in rec[0]}
        exact = {t for t in range(12) if t not in seen}
    fast_oracle = {s: v for s, v in oracle.items() if s // 4 not in exact}
    compare_records(rec, fast_oracle, 1e-6)
    check_bins(counts, edges, oracle, 1e-6)


def test_batch_exact_path_tracks():
    # a NaN point and an A3-like plane (3000 spheres per layer) force the
    # exact path; an A5-like track (1000 coincident points per sphere) stays
    # on the fused kernel; bins equal the oracle's
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=4)
    x = x.copy()
    x[off[1] + 5] = np.nan
    px, py, pz, pe = F.a3()
    keep = pz < pz.min() + 20e3               # the first 20 layers of the plane: 60000 points
    ax, ay, az, ae = F.a5_track()
    sel = slice(0, 200000)
    xs = np.concatenate([x, px[keep], ax[sel]]); ys = np.concatenate([y, py[keep], ay[sel]])
    zs = np.concatenate([z, pz[keep], az[sel]]); es = np.concatenate([e, pe[keep], ae[sel]])
    off2 = np.concatenate([off, [off[-1] + keep.sum(), off[-1] + keep.sum() + 200000]]).astype(np.int64)
    geoms = [(1.5e6, 1e3)]
    counts, _, info, edges = gpu_run(xs, ys, zs, es, off2, geoms, None, 3, 11, True, False, st.ST_F64,
                                     [("lin", 30000, 1e-1, 1e4)])
    assert info["exact_tracks"] == 2 and info["fast_tracks"] == 4
    bins, _, _ = po.run(xs, ys, zs, es, off2, [po.geometry(*geoms[0])], None, 3, 11, True, False, [edges[0]])
    np.testing.assert_array_equal(counts[0], bins[0])


def test_batch_is_deterministic_and_tuning_invariant():
    w = W.workload("cfg3")
    spec = dict(w["synth"], n_max=100000)
    x, y, z, e, off = st.synth_library(spec, 24, 50)
    results = []
    for slots, bpc, graph, ppt in [(1536, 5, True, 1), (2048, 2, False, 2), (4096, 1, True, 2), (1536, 3, False, 2),
                                   (2048, 4, True, 1), (1280, 4, False, 1)]:
        counts, _, _, _ = gpu_run(x, y, z, e, off, [(2.5e3, 10.0)], None, 8, 3, True, True, st.ST_F32,
                                  [("log", 1000, 1e-2, 1e3)], table_slots=slots, blocks_per_cu=bpc, use_graph=graph,
                                  points_per_thread=ppt)
        results.append(counts[0])
    for r in results[1:]:
        np.testing.assert_array_equal(r, results[0])
    assert results[0].sum() > 1000


def test_batch_graph_replays_match_eager():
    w = W.workload("cfg3")
    spec = dict(w["synth"], n_max=50000)
    x, y, z, e, off = st.synth_library(spec, 12, 7)
    geoms = [st.SphericalGeometry(2.5e3, 10.0)]
    lib = st.TrackLibrary(x, y, z, e, off, geoms, None, st.ST_F64)
    h = st.Histogram(1000, 1e-2, 1e3, "log")
    eager = st.BatchRun(lib, [h], 5, 9, True, True, use_graph=False)
    eager.run(); eager.sync()
    ref = h.fetch(reset=True)
    graph = st.BatchRun(lib, [h], 5, 9, True, True, use_graph=True)
    for rep in range(4):
        graph.run(); graph.sync()
        got = h.fetch(reset=True)
        np.testing.assert_array_equal(got, ref, err_msg=f"replay {rep}")
        eager.run(); eager.sync()          # eager work between replays
        np.testing.assert_array_equal(h.fetch(reset=True), ref)
    eager.free(); graph.free(); h.free(); lib.free()


@pytest.mark.parametrize("name", ["A1", "A2", "A4", "A6", "A8", "A3"])
def test_batch_reference_known_answers(name):
    fx, (kind, nb, lo, hi), nt, nos, value, count = F.MACROS[name]
    x, y, z, e = fx()
    X = np.tile(x, nt); Y = np.tile(y, nt); Z = np.tile(z, nt); E = np.tile(e, nt)
    off = np.arange(nt + 1, dtype=np.int64) * len(x)
    for precision in (st.ST_F64, st.ST_F32):
        counts, _, info, edges = gpu_run(X, Y, Z, E, off, [F.geometry_of(name)], None, nos, 1234, False, False,
                                         precision, [(kind, nb, lo, hi)])
        np.testing.assert_array_equal(counts[0].astype(np.float64), F.expected_bins(edges[0], value, count))


@pytest.mark.parametrize("precision,diameter", [(st.ST_F32, 10.0), (st.ST_F32, 2.0), (st.ST_F64, 10.0)])
def test_batch_points_on_decision_boundaries(precision, diameter):
    """Points placed within a few ulps of voxel faces, half-way planes and
    sphere surfaces *after* each oversample's shift: the FP32 fast path must
    route every such decision to the exact FP64 recheck, so keys equal the
    oracle's exactly."""
    rng = np.random.default_rng(int(diameter) + precision)
    g = po.geometry(2.5e3, diameter)
    H, D, r = g.half_length, g.diameter, g.radius
    n_tracks, n_os, per_os = 4, 3, 6000
    xs, ys, zs, es, off = [], [], [], [], [0]
    for t in range(n_tracks):
        sh = po.shift_table(W.SEED, t, n_os, True, True, g)
        px, py, pz = [], [], []
        for os in range(n_os):
            sx, sy, sz = sh[os]
            m = per_os
            k = rng.integers(0, g.n_linear, (3, m))
            base = g.gso + D * k                       # sphere centres (shifted frame)
            kind = rng.integers(0, 4, (3, m))          # 0 centre-ish, 1 half-way, 2 +surface, 3 -surface
            pos = np.where(kind == 1, base + 0.5 * D, np.where(kind == 2, base + r, np.where(kind == 3, base - r,
                                                                                           base + rng.normal(0, 0.2 * D, (3, m)))))
            face = rng.random((3, m)) < 0.05           # some on the voxel faces
            pos = np.where(face, np.where(rng.random((3, m)) < 0.5, -H, H), pos)
            pos = pos + rng.integers(-3, 4, (3, m)) * np.spacing(np.abs(pos) + 1.0)
            # back to the unshifted frame (z also carries the +H' of the z shift)
            px.append(pos[0] - sx)
            py.append(pos[1] - sy)
            pz.append(pos[2] - sz - H)
        x = np.concatenate(px); y = np.concatenate(py); z = np.concatenate(pz)
        xs.append(x); ys.append(y); zs.append(z); es.append(rng.uniform(1, 1000, len(x)))
        off.append(off[-1] + len(x))
    x, y, z, e = (np.concatenate(a) for a in (xs, ys, zs, es))
    off = np.array(off, np.int64)
    oracle = oracle_segments(x, y, z, e, off, [g], None, n_os, W.SEED, True, True)
    counts, rec, info, edges = gpu_run(x, y, z, e, off, [(2.5e3, diameter)], None, n_os, W.SEED, True, True,
                                       precision, [("log", 1000, 1e-2, 1e3)], record=400000)
    assert info["exact_tracks"] == 0
    n, _ = compare_records(rec, oracle, 1e-6 if precision == st.ST_F32 else 1e-12)
    assert n > 1000
    check_bins(counts, edges, oracle, 1e-6 if precision == st.ST_F32 else 1e-12)
    # run statistics: in-sphere points equal the oracle's pair count; the
    # points placed on boundaries are rechecked from FP64 (ST_F32 only)
    n_pairs = sum(len(pk) for t in range(n_tracks)
                  for pk, _, _, _ in po.track_segments(x[off[t]:off[t + 1]], y[off[t]:off[t + 1]],
                                                       z[off[t]:off[t + 1]], e[off[t]:off[t + 1]], g, W.SEED, t,
                                                       n_os, True, True))
    assert info["points_in_sphere"] == n_pairs
    assert n_pairs <= info["points_scored"] <= n_tracks * n_os * n_os * per_os
    if precision == st.ST_F32:
        assert info["fp64_rechecks"] > 0.2 * n_tracks * n_os * per_os
    else:
        assert info["fp64_rechecks"] == 0


@pytest.mark.parametrize("precision", [st.ST_F64, st.ST_F32])
def test_batch_ragged_library_with_empty_and_tiny_tracks(precision):
    """Empty tracks, one-point tracks and tracks of 255/256/257 points (block
    edges) mixed with ordinary ones: counts and records equal the oracle's."""
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=6)
    rng = np.random.default_rng(5)
    pieces = []
    for t in range(6):
        b, en = off[t], off[t + 1]
        pieces.append((x[b:en], y[b:en], z[b:en], e[b:en]))
    sizes = [0, 1, 255, 256, 257, 0, 513]
    orig = list(pieces)
    for k, n in enumerate(sizes):
        src = orig[k % 6]
        sel = rng.choice(len(src[0]), size=n, replace=False) if n else np.zeros(0, np.int64)
        pieces.insert(2 * k, tuple(a[sel] for a in src))
    xs, ys, zs, es = (np.concatenate([p[i] for p in pieces]) for i in range(4))
    offs = np.concatenate([[0], np.cumsum([len(p[0]) for p in pieces])]).astype(np.int64)
    geoms = [po.geometry(2.5e3, 10.0)]
    oracle = oracle_segments(xs, ys, zs, es, offs, geoms, None, 7, 21, True, True)
    counts, rec, info, edges = gpu_run(xs, ys, zs, es, offs, [(2.5e3, 10.0)], None, 7, 21, True, True, precision,
                                       [("lin", 3000, 0.01, 300.0)], record=400000)
    assert info["n_tracks"] == len(pieces) and info["exact_tracks"] == 0
    rtol = 1e-12 if precision == st.ST_F64 else 1e-6
    n, _ = compare_records(rec, oracle, rtol)
    assert n > 100
    check_bins(counts, edges, oracle, rtol)


def _small_mixed_workload():
    w = W.workload("cfg5")
    w["chunk"] = 2
    w["synth"]["n_max"] = 30000
    return w


@pytest.mark.parametrize("precision", [st.ST_F64, st.ST_F32])
def test_batch_streamed_library_matches_create(precision):
    # begin / append (runs of at most 3 same-geometry tracks) / finish lays out
    # the same library as st_library_create over the whole arrays
    w = _small_mixed_workload()
    w["chunk"] = 4   # runs of 4 tracks per geometry, appended as 3 + 1
    geoms = W.geometries(w)
    nt, npts, chunks = W.library_chunks(w, 12, max_tracks=3)
    chunks = list(chunks)
    assert [len(c[5]) for c in chunks] == [3, 1, 3, 1, 3, 1]
    x, y, z, e = (np.concatenate([c[k] for c in chunks]) for k in range(4))
    counts = np.concatenate([np.diff(c[4]) for c in chunks])
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    gt = np.concatenate([c[5] for c in chunks])
    assert nt == 12 and npts == len(x) == off[-1]
    hists = [("lin", 3000, 0.01, 300.0), ("log", 1000, 1e-2, 1e3)]
    res = []
    for streamed in (False, True):
        if streamed:
            lib = st.TrackLibrary.streamed(nt, npts, iter(chunks), geoms, precision)
        else:
            lib = st.TrackLibrary(x, y, z, e, off, geoms, gt, precision)
        hs = [st.Histogram(nb, lo, hi, kind) for kind, nb, lo, hi in hists]
        b = st.BatchRun(lib, hs, 4, 5, True, True)
        b.run()
        b.sync()
        info = lib.info()
        res.append(([h.fetch() for h in hs], b.stats(), info["fast_tracks"], info["n_points"]))
        b.free()
        for h in hs:
            h.free()
        lib.free()
    (c0, s0, f0, n0), (c1, s1, f1, n1) = res
    for a, b2 in zip(c0, c1):
        np.testing.assert_array_equal(a, b2)
    assert s0 == s1 and f0 == f1 and n0 == n1 == npts
    assert sum(a.sum() for a in c0) > 0


def test_library_streamed_misuse_is_rejected():
    w = _small_mixed_workload()
    geoms = W.geometries(w)
    nt, npts, chunks = W.library_chunks(w, 4, max_tracks=2)
    chunks = list(chunks)
    first = chunks[0]
    n_first = len(first[0])
    # more tracks than begin sized
    with pytest.raises(st.SuperTrackError, match="exceed"):
        st.TrackLibrary.streamed(1, npts, iter([first]), geoms, st.ST_F32)
    # more points than begin sized
    with pytest.raises(st.SuperTrackError, match="exceed"):
        st.TrackLibrary.streamed(nt, n_first - 1, iter([first]), geoms, st.ST_F32)
    # finish before every track was appended
    with pytest.raises(st.SuperTrackError, match="appended"):
        st.TrackLibrary.streamed(nt, npts, iter(chunks[:1]), geoms, st.ST_F32)
    # a batch over a library that was not finished
    import ctypes as C
    from supertrack_amd import _capi
    h = C.c_void_p()
    garr = (_capi.Geometry * len(geoms))(*[g.c for g in geoms])
    st.call("st_library_begin", C.c_int64(nt), C.c_int64(npts), garr, len(geoms), st.ST_F32, C.byref(h))
    lib = st.TrackLibrary.__new__(st.TrackLibrary)
    lib.h, lib.geometries = h, geoms
    hist = st.Histogram(100, 0.1, 10.0, "lin")
    with pytest.raises(st.SuperTrackError, match="not finished"):
        st.BatchRun(lib, [hist], 2, 1, True, True)
    hist.free()
    lib.free()


def test_batch_empty_library():
    geoms = [st.SphericalGeometry(2.5e3, 10.0)]
    z0 = np.zeros(0)
    lib = st.TrackLibrary(z0, z0, z0, z0, np.zeros(1, np.int64), geoms, None, st.ST_F32)
    h = st.Histogram(100, 0.1, 10.0, "lin")
    b = st.BatchRun(lib, [h], 10, 1, True, True)
    b.run()
    b.sync()
    assert h.fetch().sum() == 0
    assert b.stats()["points_scored"] == 0
    b.free(); h.free(); lib.free()


@pytest.mark.parametrize("precision", [st.ST_F64, st.ST_F32])
@pytest.mark.parametrize("H,D", [(1.0, 10.0), (5e3, 0.04)])
def test_batch_degenerate_and_fine_lattices(precision, H, D):
    """N = 1 (sphere larger than the voxel) and N = 250000 (keys above 2^53)."""
    g = po.geometry(H, D)
    rng = np.random.default_rng(int(H) + precision)
    Hp = g.half_length
    xs, ys, zs, es, off = [], [], [], [], [0]
    for t in range(5):
        n = 5000
        x = rng.uniform(-1.2 * Hp, 1.2 * Hp, n)
        y = rng.uniform(-1.2 * Hp, 1.2 * Hp, n)
        z = np.sort(rng.uniform(-1.2 * Hp, 1.2 * Hp, n)) - Hp
        if D < 1:
            x = np.round(x / D) * D + rng.normal(0, 0.2 * D, n)
        xs.append(x); ys.append(y); zs.append(z); es.append(rng.uniform(1, 1000, n))
        off.append(off[-1] + n)
    x, y, z, e = (np.concatenate(a) for a in (xs, ys, zs, es))
    off = np.array(off, np.int64)
    oracle = oracle_segments(x, y, z, e, off, [g], None, 4, 13, True, True)
    counts, rec, info, edges = gpu_run(x, y, z, e, off, [(H, D)], None, 4, 13, True, True, precision,
                                       [("log", 1000, 1e-2, 1e5)], record=400000)
    rtol = 1e-12 if precision == st.ST_F64 else 1e-6
    if info["exact_tracks"] == 0:
        n, _ = compare_records(rec, oracle, rtol)
        assert n > 0
    check_bins(counts, edges, oracle, rtol)


def test_batch_full_size_library_invariants():
    """The full 4096-track, 2 nm FP32 library (5.98e8 points, 1000 oversamples):
    size-independent properties at the real size -- a pass is deterministic,
    four launch shapes (every table size one point per thread uses by
    default, and 2048 slots) give identical histograms and statistics, and the
    statistics are consistent (in-sphere <= scored <= deposits); every pass
    conserves energy words and table entries."""
    w = W.workload("cfg4")
    x, y, z, e, off, _ = W.make_library(w)
    lib = st.TrackLibrary(x, y, z, e, off, W.geometries(w), None, st.ST_F32)
    del x, y, z, e
    deposits = int(off[-1]) * w["n_oversamples"]
    results = []
    # the default (twice: determinism), the six- and five-wave tables, and 2048
    for slots, bpc in [(1280, 7), (1280, 7), (1472, 6), (1536, 5), (2048, 4)]:
        hs = [st.Histogram(nb, lo, hi, kind) for kind, nb, lo, hi in w["hists"]]
        b = st.BatchRun(lib, hs, w["n_oversamples"], W.SEED, True, True, use_graph=True, table_slots=slots,
                        blocks_per_cu=bpc, points_per_thread=1)
        p = b.params()
        assert (p["table_slots"], p["blocks_per_cu"], p["points_per_thread"]) == (slots, bpc, 1), p
        b.run()
        b.sync()
        bal = b.balance()
        assert bal["energy_in"] == bal["energy_out"] and bal["entries_created"] == bal["entries_retired"], bal
        results.append(([h.fetch() for h in hs], b.stats(), bal))
        b.free()
        for h in hs:
            h.free()
    lib.free()
    c0, s0, k0 = results[0]
    for c, s, k in results[1:]:
        for a, b2 in zip(c0, c):
            np.testing.assert_array_equal(a, b2)
        assert s == s0
        # the energy words are per point; the entry counts are not compared
        # across runs: a flush survivor may merge into an entry that a new
        # point made first (one more entry created and retired), depending on
        # timing
        assert k["energy_in"] == k0["energy_in"]
    # every binned count is a retired entry (entries outside the histogram are retired unbinned)
    assert sum(int(a.sum()) for a in c0) <= k0["entries_retired"] <= s0["points_in_sphere"]
    assert 0 < s0["points_in_sphere"] <= s0["points_scored"] <= deposits
    assert s0["fp64_rechecks"] < s0["points_scored"] // 1000
    total = sum(int(a.sum()) for a in c0)
    assert 0 < total <= s0["points_in_sphere"]


def test_batch_fp64_whole_cfg3_against_oracle():
    """The whole 4096-track, 100-oversample FP64 workload (5.98e10 deposits)
    against the oracle's run driver, bin for bin.  A count may differ only for
    a sum within 1e-12 of an edge (about 0.02 such sums are expected in the
    whole run), so at most one sphere may move to a neighbouring bin."""
    w = W.workload("cfg3")
    x, y, z, e, off, _ = W.make_library(w)
    lib = st.TrackLibrary(x, y, z, e, off, W.geometries(w), None, st.ST_F64)
    hs = [st.Histogram(nb, lo, hi, kind) for kind, nb, lo, hi in w["hists"]]
    b = st.BatchRun(lib, hs, w["n_oversamples"], W.SEED, True, True, use_graph=True)
    b.run()
    b.sync()
    got = [h.fetch() for h in hs]
    edges = [h.edges.copy() for h in hs]
    b.free()
    for h in hs:
        h.free()
    lib.free()
    want, secs, threads = po.run(x, y, z, e, off, [po.geometry(*w["geometries"][0])], None, w["n_oversamples"],
                                 W.SEED, True, True, edges)
    diff = sum(int(np.abs(g.astype(np.int64) - v.astype(np.int64)).sum()) for g, v in zip(got, want))
    total = sum(int(v.sum()) for v in want)
    print(f"\nWHOLE cfg3 F64: {total} counts, {diff} differ (oracle {secs:.1f} s on {threads} threads)",
          flush=True)
    assert total > 10**8
    assert diff <= 2


@pytest.mark.parametrize("name,stride", [("cfg4", 10), ("cfg5", 96)])
def test_batch_bench_sample_against_oracle(name, stride):
    """The sample that the benchmark's CPU baseline times -- every stride-th
    track of the full library with all 1000 oversamples, every geometry and
    histogram -- scored by the batched FP32 kernel and by the oracle's run
    driver.  Every count is equal except where a sphere's oracle sum lies
    within 1e-6 of an edge (the FP32 energies' error bound): each such sum
    may move one count to the neighbouring bin, changing two counts."""
    w = W.workload(name)
    n_sample = (w["n_tracks"] + stride - 1) // stride
    x, y, z, e, off, gt = W.make_library(w, n_tracks=n_sample, stride=stride)
    mixed = len(w["geometries"]) > 1
    print(f"\nbench sample {name}: {n_sample} tracks, {int(off[-1])} points x {w['n_oversamples']} oversamples",
          flush=True)
    counts, _, info, edges = gpu_run(x, y, z, e, off, w["geometries"], gt if mixed else None, w["n_oversamples"],
                                     W.SEED, True, True, w["precision"], w["hists"], use_graph=True)
    assert info["exact_tracks"] == 0
    want, secs, threads, near = po.run(x, y, z, e, off, [po.geometry(*g) for g in w["geometries"]],
                                       gt if mixed else None, w["n_oversamples"], W.SEED, True, True, edges,
                                       near_rtol=1e-6)
    total = 0
    for h, (g, v) in enumerate(zip(counts, want)):
        diff = int(np.abs(g.astype(np.int64) - v.astype(np.int64)).sum())
        total += int(v.sum())
        print(f"PARITY bench sample {name} hist {h}: {int(v.sum())} counts, {diff} differ, {near[h]} sums within "
              f"1e-6 of an edge (oracle {secs:.1f} s on {threads} threads)", flush=True)
        assert diff <= 2 * near[h], (h, diff, near[h])
    assert total > 10**7


def test_batch_histogram_copies_are_invariant():
    """1, 2 and 4 LDS copies of the privatised histograms per workgroup give
    identical counts (one log histogram; and lin 3000 + log 1000 together,
    where four copies do not fit next to the table and are refused)."""
    w = W.workload("cfg3")
    spec = dict(w["synth"], n_max=80000)
    x, y, z, e, off = st.synth_library(spec, 20, 300)
    for hists, copies in (([("log", 1000, 1e-2, 1e3)], (1, 2, 4)),
                          ([("lin", 3000, 0.01, 300.0), ("log", 1000, 1e-2, 1e3)], (1, 2))):
        results = []
        for c in copies:
            counts, _, info, _ = gpu_run(x, y, z, e, off, [(2.5e3, 10.0)], None, 6, 8, True, True, st.ST_F32,
                                         hists, hist_copies=c)
            results.append(counts)
        for r in results[1:]:
            for a, b in zip(r, results[0]):
                np.testing.assert_array_equal(a, b)
        assert sum(int(a.sum()) for a in results[0]) > 1000
    with pytest.raises(st.SuperTrackError, match="copies"):
        gpu_run(x, y, z, e, off, [(2.5e3, 10.0)], None, 2, 8, True, True, st.ST_F32,
                [("lin", 3000, 0.01, 300.0), ("log", 1000, 1e-2, 1e3)], hist_copies=4)


@pytest.mark.parametrize("precision", [st.ST_F64, st.ST_F32])
@pytest.mark.parametrize("shift", [False, True])
def test_batch_wide_tracks_scored_in_slabs(precision, shift):
    """Tracks spread over the whole voxel of a 2000^3 lattice (1 nm spheres,
    half-length 1 um): a segment's box of reachable spheres exceeds the 2^32 - 1
    keys of the sphere table (about 2000^3 without shifts), so the kernel
    scores it in slabs of layers.  Keys and sums equal the oracle's."""
    rng = np.random.default_rng(11 + precision + 2 * shift)
    H = 1.0e3
    g = po.geometry(H, 1.0)
    assert g.n_linear == 2000
    n_tracks, per_track, n_os = 3, 20000, 3
    x = rng.uniform(-H, H, n_tracks * per_track)
    y = rng.uniform(-H, H, n_tracks * per_track)
    z = rng.uniform(-2 * H, 0.0, n_tracks * per_track)   # z + H' inside the voxel without shifts
    e = rng.uniform(1.0, 1000.0, n_tracks * per_track)
    off = np.arange(n_tracks + 1, dtype=np.int64) * per_track
    # the box of one unshifted segment: more spheres than 32-bit keys
    assert (2 * H / g.diameter) ** 3 > 2**32
    oracle = oracle_segments(x, y, z, e, off, [g], None, n_os, W.SEED, shift, shift)
    n_rec = sum(len(k) for k, _ in oracle.values())
    counts, rec, info, edges = gpu_run(x, y, z, e, off, [(H, 1.0)], None, n_os, W.SEED, shift, shift, precision,
                                       [("log", 1000, 1e-2, 1e3)], record=n_rec + 1024)
    assert info["exact_tracks"] == 0
    n, _ = compare_records(rec, oracle, 1e-6 if precision == st.ST_F32 else 1e-12)
    assert n > 10000
    check_bins(counts, edges, oracle, 1e-6 if precision == st.ST_F32 else 1e-12)
    n_pairs = sum(len(pk) for t in range(n_tracks)
                  for pk, _, _, _ in po.track_segments(x[off[t]:off[t + 1]], y[off[t]:off[t + 1]],
                                                       z[off[t]:off[t + 1]], e[off[t]:off[t + 1]], g, W.SEED, t,
                                                       n_os, shift, shift))
    assert info["points_in_sphere"] == n_pairs


def _face_alias_track(partner):
    """One point within an ulp of the +x voxel face, on the lattice line of
    sphere row (ky, kz): its x index rounds to N, and the reference's key
    kx + N ky + N^2 kz then names sphere (0, ky + 1, kz).  With `partner`, a
    second point sits at the centre of that sphere (so the two share one sum
    in the reference), plus an ordinary point in between."""
    g = po.geometry(2.5e3, 10.0)
    N, H, gso, D = g.n_linear, g.half_length, g.gso, g.diameter
    ky, kz = 200, 300
    zc = (gso + D * kz) - H               # z + H' = the layer's centre (no shift)
    x = [np.nextafter(H, 0.0)]
    y = [gso + D * ky]
    z = [zc]
    if partner:
        x += [gso, 0.3]
        y += [gso + D * (ky + 1), gso + D * ky + 1.0]
        z += [zc, zc + 1.0]
    x, y, z = (np.array(a, np.float64) for a in (x, y, z))
    e = np.linspace(10.0, 20.0, len(x))
    # the oracle's keys: the face point's key equals that of sphere (0, ky + 1, kz)
    alias_key = (ky + 1) * N + kz * N * N
    pk, _, keys, sums = po.track_segments(x, y, z, e, g, W.SEED, 0, 1, False, False)[0]
    assert int(np.count_nonzero(pk == alias_key)) == (2 if partner else 1)
    return g, x, y, z, e, alias_key, keys, sums


@pytest.mark.parametrize("precision", [st.ST_F64, st.ST_F32])
def test_batch_face_alias_merges_like_reference(precision):
    """A point whose x index rounds to N shares the reference's key with sphere
    (0, ky + 1, kz); when that sphere lies in the segment's key box, the two
    points accumulate into one entry, with the oracle's key and sum."""
    g, x, y, z, e, alias_key, keys, sums = _face_alias_track(True)
    off = np.array([0, len(x)], np.int64)
    oracle = {0: (keys, sums)}
    counts, rec, info, edges = gpu_run(x, y, z, e, off, [(2.5e3, 10.0)], None, 1, W.SEED, False, False, precision,
                                       [("log", 1000, 1e-2, 1e3)], record=64)
    n, _ = compare_records(rec, oracle, 1e-12 if precision == st.ST_F64 else 1e-6)
    assert n == len(keys) and alias_key in set(int(k) for k in rec[1])
    check_bins(counts, edges, oracle, 1e-12 if precision == st.ST_F64 else 1e-6)


@pytest.mark.parametrize("precision", [st.ST_F64, st.ST_F32])
def test_batch_face_alias_outside_key_box_is_reported(precision):
    """The same face point alone: sphere (0, ky + 1, kz) lies outside the
    segment's key box, which the 32-bit table keys cannot express; the pass
    reports it (error code 4) instead of scoring it under another key."""
    g, x, y, z, e, alias_key, keys, sums = _face_alias_track(False)
    off = np.array([0, len(x)], np.int64)
    with pytest.raises(st.SuperTrackError, match="code 4"):
        gpu_run(x, y, z, e, off, [(2.5e3, 10.0)], None, 1, W.SEED, False, False, precision,
                [("log", 1000, 1e-2, 1e3)])
