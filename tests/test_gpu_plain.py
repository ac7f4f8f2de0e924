"""Parity of the plain (per-track drop-in) HIP path with the FP64 oracle:
pair lists, per-sphere sums and histogram counts bit-identical, plus the
reference test suite's known answers."""
import numpy as np
import pytest

import fixtures as F
import pyoracle as po
import supertrack_amd as st
from supertrack_amd import workloads as W

pytestmark = pytest.mark.gpu


def gpu_track_pairs(x, y, z, e, H, D, seed, track_index, n_os, shift_xy=True, shift_z=False):
    """Runs the drop-in method on one track; returns the pair lists per oversample."""
    g = st.SphericalGeometry(H, D)
    m = st.VoxelConstrainedSphereMethod(g, shift_xy, shift_z)
    trk = st.Track(x, y, z, e)
    pairs = st.VolumeEdepPair(len(x))
    m.allocate_track_process(trk, seed, track_index, n_os)
    out = []
    for _ in range(n_os):
        m.process_track(trk, pairs)
        out.append(pairs.download())
    m.free_track_process()
    zp = trk.download()[2]
    m.free()
    return out, zp


def assert_pairs_equal(x, y, z, e, H, D, seed, ti, n_os, shift_xy=True, shift_z=False):
    got, zp = gpu_track_pairs(x, y, z, e, H, D, seed, ti, n_os, shift_xy, shift_z)
    g = po.geometry(H, D)
    want = po.track_segments(x, y, z, e, g, seed, ti, n_os, shift_xy, shift_z)
    # the method moves z by +H' in place, like ShiftTrackZOrigin
    np.testing.assert_array_equal(zp, np.asarray(z) + g.half_length)
    total = 0
    for k, ((gk, gy), (wk, wy, _, _)) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(gk, wk, err_msg=f"keys, oversample {k}")
        np.testing.assert_array_equal(gy, wy, err_msg=f"lineal energies, oversample {k}")
        total += len(gk)
    return total


def test_pairs_cfg1_track():
    w = W.workload("cfg1")
    x, y, z, e, off, _ = W.make_library(w)
    n = assert_pairs_equal(x, y, z, e, 2.5e3, 10.0, W.SEED, 0, 10, True, True)
    assert n > 500


def test_pairs_cfg2_tracks_reference_mode_and_z_shift():
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=6)
    for t in range(6):
        b, en = off[t], off[t + 1]
        assert_pairs_equal(x[b:en], y[b:en], z[b:en], e[b:en], 2.5e3, 10.0, 99, t, 4, True, t % 2 == 1)


def test_pairs_a8_keys():
    x, y, z, e = F.a8()
    got, _ = gpu_track_pairs(x, y, z, e, 1.5e6, 10.0, 1, 0, 1, False, False)
    assert [int(k) for k in got[0][0]] == F.A8_KEYS


def test_pairs_edge_cases():
    g = po.geometry(4.0, 1.0)
    H = g.half_length
    # on faces, half-way planes, sphere surfaces, NaN / Inf, far outside
    x = np.array([H, -H, 0.5, 0.0, 1.0, 2.0, np.nan, np.inf, 1e300, 0.5, 0.25, -3.5 + 1e-12])
    y = np.array([0.5, 0.5, H, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.0, 0.5, 0.5])
    z = np.array([0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, np.nan, 0.5]) - H
    e = np.arange(1, len(x) + 1, dtype=np.float64)
    assert_pairs_equal(x, y, z, e, 4.0, 1.0, 3, 0, 1, False, False)


def test_pairs_empty_and_all_outside():
    got, _ = gpu_track_pairs(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), 2.5e3, 10.0, 1, 0, 2)
    assert all(len(k) == 0 for k, _ in got)
    x = np.full(3000, 1e6)
    got, _ = gpu_track_pairs(x, x, x, np.ones(3000), 2.5e3, 10.0, 1, 0, 2)
    assert all(len(k) == 0 for k, _ in got)


def test_pairs_random_near_boundaries():
    """Points placed within a few ulps of half-way planes and sphere surfaces."""
    rng = np.random.default_rng(11)
    g = po.geometry(2.5e3, 10.0)
    n = 50000
    # lattice positions of half-way planes and of sphere surfaces along x
    k = rng.integers(0, 500, n)
    base = g.gso + g.diameter * k
    sel = rng.integers(0, 3, n)
    xs = np.where(sel == 0, base + 0.5 * g.diameter, np.where(sel == 1, base + g.radius, base - g.radius))
    xs = xs + rng.integers(-4, 5, n) * np.spacing(np.abs(xs) + 1.0)
    y = g.gso + g.diameter * rng.integers(0, 500, n) + rng.normal(0, 1e-3, n)
    z = g.gso + g.diameter * rng.integers(0, 500, n) - g.half_length
    e = rng.uniform(1, 1000, n)
    assert_pairs_equal(xs, y, z, e, 2.5e3, 10.0, 5, 0, 1, False, False)


def test_pairs_track_of_more_than_64_tiles():
    # 150,000 points = 147 compaction tiles of 1024: the look-back walks more
    # than one 64-tile window
    rng = np.random.default_rng(150)
    n = 150000
    H = po.geometry(2.5e3, 10.0).half_length
    x = rng.uniform(-H, H, n)
    y = rng.uniform(-H, H, n)
    z = rng.uniform(-2 * H, 0.0, n)
    e = rng.uniform(1.0, 100.0, n)
    assert assert_pairs_equal(x, y, z, e, 2.5e3, 10.0, 21, 3, 3, True, True) > 100000


def test_pairs_large_track():
    w = W.workload("cfg3")
    spec = dict(w["synth"], entry_half_length=2500.0, voxel_half_length=2500.0)
    x, y, z, e, off = st.synth_library(spec, 1, 17)
    assert len(x) > 20000
    assert_pairs_equal(x, y, z, e, 2.5e3, 10.0, 4, 17, 3, True, True)


# ---------------------------------------------------------------------------
# Histogram: sort -> reduce -> bin on given pairs
# ---------------------------------------------------------------------------
def hist_from_pairs(kind, nb, lo, hi, keys, ys, capacity=None):
    h = st.Histogram(nb, lo, hi, kind)
    p = st.VolumeEdepPair(capacity or max(len(keys), 1))
    h.allocate_track_process(p)
    p.upload(keys, ys)
    h.sort_reduce_and_add_to_histogram(p)
    h.free_track_process()
    out = h.counts()
    edges = h.edges.copy()
    h.free()
    p.free()
    return edges, out


@pytest.mark.parametrize("capacity", [None, 20000])
def test_histogram_tests_known_answers(capacity):
    # tst/tests/HistogramTests.cu; capacity 20000 takes the multi-workgroup path
    from test_oracle import HIST_CASES
    for kind, nb, lo, hi, keys, v, binned, count in HIST_CASES:
        edges, got = hist_from_pairs(kind, nb, lo, hi, keys.astype(np.uint64), np.full(len(keys), v), capacity)
        np.testing.assert_array_equal(edges, po.edges(kind, lo, hi, nb))
        np.testing.assert_array_equal(got, F.expected_bins(edges, binned, count))


@pytest.mark.parametrize("m,keyspace,capacity", [(1, 10, None), (5000, 1 << 20, None), (6144, 7, None),
                                                 (6144, 1 << 40, None), (6145, 1 << 27, None),
                                                 (250000, 1 << 27, None), (300000, 1 << 60, None),
                                                 (4000, 1 << 27, 100000), (0, 10, 100000),
                                                 (100000, 1 << 35, 200000), (131072, 1 << 20, None),
                                                 (20000, 3, None), (12288, 1 << 20, None), (12289, 1 << 40, None),
                                                 (12000, 1, None), (9000, 1 << 62, 12288),
                                                 (3000, 2, None), (4096, 1 << 50, None), (4097, 4, None),
                                                 (5120, 5, None), (5121, 3, None), (1536, 1 << 30, None),
                                                 (1537, 5, None), (1000, 2, None)])
def test_histogram_random_pairs_match_oracle(m, keyspace, capacity):
    rng = np.random.default_rng(m + 3)
    keys = rng.integers(0, keyspace, m, dtype=np.uint64)
    # many repeats so that runs are long
    keys = keys[rng.integers(0, max(m, 1), m)] if m else keys
    ys = rng.lognormal(0, 1, m)
    edges, got = hist_from_pairs("log", 1000, 1e-2, 1e3, keys, ys, capacity)
    _, sums = po.reduce_pairs(keys, ys)
    want = po.histogram_counts(edges, sums)
    np.testing.assert_array_equal(got, want.astype(np.float64))


@pytest.mark.parametrize("nb", [2000, 4095, 4096, 4097, 20000])
def test_histogram_many_bins_match_oracle(nb):
    # up to 4096 bins are counted in LDS per launch, more go to global counters;
    # edges of up to 4095 bins are staged in LDS when they fit beside the pairs
    rng = np.random.default_rng(nb)
    m = 3000
    keys = rng.integers(0, 200, m).astype(np.uint64)
    ys = rng.uniform(0.0, 2.0, m)
    edges, got = hist_from_pairs("lin", nb, 0.0, 60.0, keys, ys)
    _, sums = po.reduce_pairs(keys, ys)
    want = po.histogram_counts(edges, sums)
    assert want.sum() == len(np.unique(keys))
    np.testing.assert_array_equal(got, want.astype(np.float64))


# ---------------------------------------------------------------------------
# whole per-track loop vs the oracle's run driver
# ---------------------------------------------------------------------------
def test_plain_loop_bins_match_oracle_cfg2_subset():
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=12)
    kind, nb, lo, hi = w["hists"][0]
    h = st.Histogram(nb, lo, hi, kind)
    g = st.SphericalGeometry(2.5e3, 10.0)
    st.run_plain(x, y, z, e, off, g, h, 20, W.SEED, True, True)
    bins, _, _ = po.run(x, y, z, e, off, [po.geometry(2.5e3, 10.0)], None, 20, W.SEED, True, True,
                        [po.edges(kind, lo, hi, nb)])
    np.testing.assert_array_equal(h.counts(), bins[0].astype(np.float64))
    assert bins[0].sum() > 1000


def test_plain_loop_bins_match_oracle_cfg2_full():
    # the whole 256-track, 100-oversample workload, bin for bin
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w)
    kind, nb, lo, hi = w["hists"][0]
    h = st.Histogram(nb, lo, hi, kind)
    g = st.SphericalGeometry(2.5e3, 10.0)
    n_os = w["n_oversamples"]
    st.run_plain(x, y, z, e, off, g, h, n_os, W.SEED, True, True)
    bins, _, _ = po.run(x, y, z, e, off, [po.geometry(2.5e3, 10.0)], None, n_os, W.SEED, True, True,
                        [po.edges(kind, lo, hi, nb)])
    np.testing.assert_array_equal(h.counts(), bins[0].astype(np.float64))
    assert len(off) - 1 == 256 and bins[0].sum() > 1e6


def test_plain_loop_is_deterministic():
    w = W.workload("cfg1")
    x, y, z, e, off, _ = W.make_library(w)
    res = []
    for _ in range(2):
        h = st.Histogram(3000, 0.01, 300.0, "lin")
        st.run_plain(x, y, z, e, off, st.SphericalGeometry(2.5e3, 10.0), h, 10, W.SEED, True, True)
        res.append(h.counts())
    np.testing.assert_array_equal(res[0], res[1])


@pytest.mark.parametrize("name", ["A1", "A2", "A3", "A4", "A5", "A6", "A8"])
def test_reference_integration_known_answers(name):
    fx, (kind, nb, lo, hi), nt, nos, value, count = F.MACROS[name]
    x, y, z, e = fx()
    X = np.tile(x, nt); Y = np.tile(y, nt); Z = np.tile(z, nt); E = np.tile(e, nt)
    off = np.arange(nt + 1, dtype=np.int64) * len(x)
    h = st.Histogram(nb, lo, hi, kind)
    st.run_plain(X, Y, Z, E, off, st.SphericalGeometry(*F.geometry_of(name)), h, nos, 1234, False, False)
    np.testing.assert_array_equal(h.counts(), F.expected_bins(h.edges, value, count))


def test_reference_a7_pi_estimate():
    x, y, z, e = F.a7()
    off = np.array([0, len(x)], np.int64)
    h = st.Histogram(10000, 1, 1e4, "lin")
    st.run_plain(x, y, z, e, off, st.SphericalGeometry(1.5e6, 1e3), h, 10000, 20240601, True, False)
    mid = (h.edges[:-1] + h.edges[1:]) / 2
    pi = 6 * float(np.sum(mid * h.counts())) / (20000 * 10000)
    assert 3.05 < pi < 3.3, pi
    # and equal to the oracle's count-for-count
    bins, _, _ = po.run(x, y, z, e, off, [po.geometry(1.5e6, 1e3)], None, 10000, 20240601, True, False,
                        [po.edges("lin", 1, 1e4, 10000)])
    np.testing.assert_array_equal(h.counts(), bins[0].astype(np.float64))


def test_plain_concurrent_workers_on_streams_match_oracle():
    """Workers on their own HIP streams in separate host threads (the
    reference's forked workers): the summed histogram equals the oracle's."""
    import ctypes as C
    import threading
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=18)
    kind, nb, lo, hi = w["hists"][0]
    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    n_w = 3
    hs, errs = [], []

    def worker(k, stream):
        try:
            h = st.Histogram(nb, lo, hi, kind)
            h.set_stream(stream)
            m = st.VoxelConstrainedSphereMethod(st.SphericalGeometry(2.5e3, 10.0), True, True)
            m.set_stream(stream)
            for t in range(k, 18, n_w):
                b, en = off[t], off[t + 1]
                trk = st.Track(x[b:en], y[b:en], z[b:en], e[b:en])
                pairs = st.VolumeEdepPair(en - b)
                m.allocate_track_process(trk, W.SEED, t, 20)
                h.allocate_track_process(pairs)
                for _ in range(20):
                    m.process_track(trk, pairs)
                    h.sort_reduce_and_add_to_histogram(pairs)
                m.free_track_process()
                h.free_track_process()
            hs.append(h.counts())
            m.free()
            h.free()
        except Exception as ex:   # reported by the main thread
            errs.append(ex)

    streams = []
    for _ in range(n_w):
        s = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(s)) == 0
        streams.append(s)
    threads = [threading.Thread(target=worker, args=(k, streams[k])) for k in range(n_w)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    st.synchronize()
    for s in streams:
        hip.hipStreamDestroy(s)
    assert not errs, errs
    bins, _, _ = po.run(x, y, z, e, off, [po.geometry(2.5e3, 10.0)], None, 20, W.SEED, True, True,
                        [po.edges(kind, lo, hi, nb)])
    np.testing.assert_array_equal(sum(hs), bins[0].astype(np.float64))


@pytest.mark.parametrize("H,D", [(1.0, 10.0), (5e3, 0.04)])
def test_pairs_degenerate_and_fine_lattices(H, D):
    """The sphere-larger-than-voxel branch (N = 1) and a lattice with keys
    above 2^53 (N = 250000): pairs equal the oracle's."""
    g = po.geometry(H, D)
    rng = np.random.default_rng(int(H))
    n = 20000
    Hp = g.half_length
    x = rng.uniform(-1.2 * Hp, 1.2 * Hp, n)
    y = rng.uniform(-1.2 * Hp, 1.2 * Hp, n)
    z = rng.uniform(-1.2 * Hp, 1.2 * Hp, n) - Hp
    if D < 1:   # concentrate points so that spheres collect several
        x = np.round(x / D) * D + rng.normal(0, 0.2 * D, n)
        y = np.round(y / D) * D + rng.normal(0, 0.2 * D, n)
    e = rng.uniform(1, 1000, n)
    total = assert_pairs_equal(x, y, z, e, H, D, 8, 3, 3, True, True)
    assert total > 100
