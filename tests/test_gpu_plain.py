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


def test_num_elements_is_host_readable_after_the_track_sync():
    """VolumeEdepPair::numElements is read on the host after the track's
    synchronisation (VolumeEdepPair.cu:10-13, Histogram.cu:105): through the C
    ABI, after st_vcs_end_track the count is read through the plain pointer,
    and equals the oracle's pair count of the last oversample."""
    import ctypes as C
    from supertrack_amd import _capi
    w = W.workload("cfg1")
    x, y, z, e, off, _ = W.make_library(w)
    n = len(x)
    g = _capi.Geometry()
    st.call("st_geometry_make", C.c_double(2.5e3), C.c_double(10.0), C.byref(g))
    m = C.c_void_p()
    st.call("st_vcs_create", C.byref(g), 1, 1, 0, 0, C.byref(m))
    trk = _capi.Track()
    arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, e)]
    st.call("st_track_upload", C.byref(trk), *[a.ctypes.data_as(_capi.c_double_p) for a in arrs], C.c_int64(n))
    pairs = _capi.Pairs()
    st.call("st_pairs_alloc", C.byref(pairs), C.c_int64(n))
    count = C.cast(pairs.num_elements, C.POINTER(C.c_int32))
    assert count[0] == n                       # starts at the capacity
    n_os = 6
    want = po.track_segments(x, y, z, e, po.geometry(2.5e3, 10.0), 5, 0, n_os, True, True)
    assert sum(len(wk) for wk, _, _, _ in want) > 0
    st.call("st_vcs_begin_track", m, C.byref(trk), C.c_uint64(5), C.c_uint32(0), C.c_int32(n_os))
    for k in range(n_os):
        st.call("st_vcs_process_oversample", m, C.byref(trk), C.byref(pairs))
        if k < n_os - 1:
            st.call("st_device_synchronize")
            assert count[0] == len(want[k][0]), k
    st.call("st_vcs_end_track", m)
    assert count[0] == len(want[-1][0])
    st.lib().st_pairs_free(C.byref(pairs))
    st.lib().st_track_free(C.byref(trk))
    st.lib().st_vcs_destroy(m)


def test_process_exit_with_live_objects_is_clean(tmp_path):
    """A program that leaves every kind of library object alive (and work
    queued) at exit ends with status 0: the binding frees them and calls
    st_shutdown before the HIP runtime is torn down."""
    import subprocess
    import sys
    import os
    script = tmp_path / "live_objects.py"
    script.write_text(f"""
import sys
sys.path.insert(0, {os.path.dirname(os.path.dirname(os.path.abspath(__file__)))!r})
import numpy as np
import supertrack_amd as st
from supertrack_amd import workloads as W
w = W.workload("cfg2")
x, y, z, e, off, _ = W.make_library(w, n_tracks=4)
g = st.SphericalGeometry(2.5e3, 10.0)
h = st.Histogram(3000, 0.01, 300.0, "lin")
st.run_plain(x, y, z, e, off, g, h, 3, 1, True, True)
m = st.VoxelConstrainedSphereMethod(g, True, True)
trk = st.Track(x[:off[1]], y[:off[1]], z[:off[1]], e[:off[1]])
pairs = st.VolumeEdepPair(int(off[1]))
m.allocate_track_process(trk, 1, 0, 5)
h.allocate_track_process(pairs)
m.process_track(trk, pairs)
h.sort_reduce_and_add_to_histogram(pairs)
lib = st.TrackLibrary(x, y, z, e, off, [g], None, st.ST_F32)
hb = st.Histogram(1000, 1e-2, 1e3, "log")
b = st.BatchRun(lib, [hb], 50, 1, True, True)
b.run()
d = st.DeviceArray(1 << 20)
print("objects alive at exit", flush=True)
""")
    res = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, (res.returncode, res.stdout[-2000:], res.stderr[-2000:])
    assert "objects alive at exit" in res.stdout


@pytest.mark.parametrize("shift_z", [False, True])
def test_queued_oversamples_equal_the_per_oversample_path(shift_z):
    """Queued oversamples (one launch per track at its end) give the same bins
    as the per-oversample path and the oracle, bit for bit, on cfg2 tracks
    (including one above the 12288-point queue limit, which runs oversample
    by oversample); after the track the pairs hold the last oversample's."""
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=6)
    big = st.synth_library(dict(W.workload("cfg3")["synth"], n_min=20000, n_max=20000), 1, 3)
    x = np.concatenate([x, big[0]]); y = np.concatenate([y, big[1]])
    z = np.concatenate([z, big[2]]); e = np.concatenate([e, big[3]])
    off = np.concatenate([off, [off[-1] + len(big[0])]]).astype(np.int64)
    assert int(off[-1] - off[-2]) > 12288
    g = st.SphericalGeometry(2.5e3, 10.0)
    n_os = 9
    hq = st.Histogram(3000, 0.01, 300.0, "lin")
    hp = st.Histogram(3000, 0.01, 300.0, "lin")
    st.run_plain(x, y, z, e, off, g, hq, n_os, 7, True, shift_z, queued=True)
    st.run_plain(x, y, z, e, off, g, hp, n_os, 7, True, shift_z)
    ed = po.edges("lin", 0.01, 300.0, 3000)
    bins, _, _ = po.run(x, y, z, e, off, [po.geometry(2.5e3, 10.0)], None, n_os, 7, True, shift_z, [ed])
    np.testing.assert_array_equal(hq.counts(), bins[0].astype(np.float64))
    np.testing.assert_array_equal(hp.counts(), bins[0].astype(np.float64))
    assert bins[0].sum() > 1000
    # the pairs after a queued track: the last oversample's, in point order
    t = 2
    b, en = off[t], off[t + 1]
    m = st.VoxelConstrainedSphereMethod(g, True, shift_z)
    m.set_queued(True)
    trk = st.Track(x[b:en], y[b:en], z[b:en], e[b:en])
    pairs = st.VolumeEdepPair(int(en - b))
    h = st.Histogram(3000, 0.01, 300.0, "lin")
    m.allocate_track_process(trk, 7, t, n_os)
    h.allocate_track_process(pairs)
    for _ in range(n_os):
        m.process_track(trk, pairs)
        h.sort_reduce_and_add_to_histogram(pairs)
    m.free_track_process()
    h.free_track_process()
    want = po.track_segments(x[b:en], y[b:en], z[b:en], e[b:en], po.geometry(2.5e3, 10.0), 7, t, n_os, True,
                             shift_z)
    gk, gy = pairs.download()
    np.testing.assert_array_equal(gk, want[-1][0])
    np.testing.assert_array_equal(gy, want[-1][1])
    for o in (m, h, trk, pairs):
        o.free()


def test_queued_oversamples_uneven_consumers_and_streams():
    """Two histograms consuming different oversamples of one queued track, one
    of them on another stream; a histogram read before the track ends sees
    the pending work; pairs freed while oversamples are pending."""
    import ctypes as C
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=1)
    n = int(off[1])
    g = st.SphericalGeometry(2.5e3, 10.0)
    gp = po.geometry(2.5e3, 10.0)
    n_os = 7
    segs = po.track_segments(x, y, z, e, gp, 3, 0, n_os, True, True)
    ed_lin = po.edges("lin", 0.01, 300.0, 3000)
    ed_log = po.edges("log", 1e-2, 1e3, 1000)
    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    other = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(other)) == 0
    m = st.VoxelConstrainedSphereMethod(g, True, True)
    m.set_queued(True)
    trk = st.Track(x, y, z, e)
    pairs = st.VolumeEdepPair(n)
    ha = st.Histogram(3000, 0.01, 300.0, "lin")
    hb = st.Histogram(1000, 1e-2, 1e3, "log")
    hb.set_stream(other)
    m.allocate_track_process(trk, 3, 0, n_os)
    ha.allocate_track_process(pairs)
    hb.allocate_track_process(pairs)
    for k in range(n_os):
        m.process_track(trk, pairs)
        ha.sort_reduce_and_add_to_histogram(pairs)          # every oversample
        if k % 3 == 1:
            hb.sort_reduce_and_add_to_histogram(pairs)      # oversamples 1 and 4 only
            hb.sort_reduce_and_add_to_histogram(pairs)      # twice: counted twice, as on the per-oversample path
    # reading hb before the method's track ends runs the pending oversamples
    got_b = hb.fetch(reset=False)
    want_b = sum(2 * po.histogram_counts(ed_log, segs[k][3]) for k in (1, 4))
    np.testing.assert_array_equal(got_b, want_b)
    m.free_track_process()
    ha.free_track_process()
    want_a = sum(po.histogram_counts(ed_lin, s[3]) for s in segs)
    np.testing.assert_array_equal(ha.counts(), want_a.astype(np.float64))
    # pending oversamples, then the pairs are freed: they run first
    m.allocate_track_process(trk, 3, 1, 2)
    m.process_track(trk, pairs)
    pairs.free()
    m.free_track_process()
    for o in (m, trk, ha, hb):
        o.free()
    hip.hipStreamDestroy(other)
