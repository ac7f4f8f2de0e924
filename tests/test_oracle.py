"""Pins the FP64 oracle (oracle/) against known answers, an independent
Philox implementation, closed forms and brute force.  CPU only."""
import math

import numpy as np
import pytest

import fixtures as F
import pyoracle as po

# ---------------------------------------------------------------------------
# Philox4x32-10
# ---------------------------------------------------------------------------
# Known-answer vectors distributed with Random123 (kat_vectors, philox4x32 10).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(int(v) for v in po.philox(ctr, key)) == want


def test_philox_matches_rocrand_engine():
    rng = np.random.default_rng(7)
    rows = [tuple(int(v) for v in rng.integers(0, 2**32, 6, dtype=np.uint64)) for _ in range(300)]
    rows += [(k, 3, 0x53484654, 0, 1234, 0) for k in range(20)]
    theirs = po.rocrand_philox(rows)
    for r, t in zip(rows, theirs):
        assert tuple(int(v) for v in po.philox(r[:4], r[4:])) == t


def test_u01_is_open_closed():
    assert po.u01(0) == 2.0**-24
    assert po.u01(0xFFFFFFFF) == 1.0
    assert po.u01(0x80000000) == (0x800000 + 1) * 2.0**-24


# ---------------------------------------------------------------------------
# geometry (SphericalGeometry.cu:34-58)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("H,D,N,Hp,gso", [
    (1.5e6, 1e3, 3000, 1.5e6, -1499500.0),
    (1.5e6, 10.0, 300000, 1.5e6, -1499995.0),
    (2.5e3, 10.0, 500, 2500.0, -2495.0),
    (2.5e3, 2.0, 2500, 2500.0, -2499.0),
    (1e3, 1.0, 2000, 1000.0, -999.5),
    (5e3, 100.0, 100, 5000.0, -4950.0),
    (1.0, 0.3, 6, 0.9, -0.75),           # snapped: 6 spheres of 0.3 span 1.8
])
def test_geometry_closed_form(H, D, N, Hp, gso):
    g = po.geometry(H, D)
    assert g.n_linear == N
    assert g.half_length == pytest.approx(Hp, rel=1e-15)
    assert g.gso == pytest.approx(gso, rel=1e-15)
    assert g.length == N * D
    assert g.radius == D / 2


def test_geometry_sphere_larger_than_voxel():
    g = po.geometry(1.0, 5.0)   # r = 2.5 > H' = 0
    assert g.n_linear == 1 and g.half_length == 2.5 and g.length == 5.0 and g.gso == 0.0


# ---------------------------------------------------------------------------
# bin edges (utils.cc:8-31) and HistogramRange bin selection
# ---------------------------------------------------------------------------
def test_edges_carry_float_truncation():
    e = po.edges("lin", 0.01, 300.0, 3000)
    assert e[0] == 0.0099999997764825821
    assert e[1] == 0.10999666340649128
    assert e[3000] == 299.99999088980258
    el = po.edges("log", 1e-2, 1e3, 1000)
    assert el[1000] == 999.99974266609581


def test_edges_float_arithmetic_restated():
    # The same float/double mixture written with numpy's float32 (an
    # independent evaluation): width is a float quotient widened to double.
    for lo, hi, nb in [(0.01, 300.0, 3000), (0.1, 1e7, 30000), (1.0, 1e4, 10000), (0.0, 500.0, 500)]:
        lo32, hi32 = np.float32(lo), np.float32(hi)
        w = np.float64((hi32 - lo32) / np.float32(nb))
        want = np.float64(lo32) + np.arange(nb + 1, dtype=np.float64) * w
        np.testing.assert_array_equal(po.edges("lin", lo, hi, nb), want)
    for lo, hi, nb in [(1e-2, 1e3, 1000), (0.1, 1e7, 30000)]:
        bo = np.float32(math.log10(np.float32(lo)))
        to = np.float32(math.log10(np.float32(hi)))
        w = np.float64((to - bo) / np.float32(nb))
        want = np.array([10.0 ** (np.float64(bo) + i * w) for i in range(nb + 1)])
        np.testing.assert_allclose(po.edges("log", lo, hi, nb), want, rtol=4e-16)


def test_log_edges_from_zero_are_nan_and_count_nothing():
    e = po.edges("log", 0.0, 500.0, 500)
    assert np.all(np.isnan(e))
    assert po.bin_index(e, 5.0) == -1


def test_bin_matches_upper_bound():
    e = po.edges("lin", 0.01, 300.0, 3000)
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.uniform(-1, 310, 3000), e, np.nextafter(e, -np.inf), np.nextafter(e, np.inf),
                           [np.nan, np.inf, -np.inf, 0.0, 300.0]])
    for v in vals:
        want = -1
        if e[0] <= v < e[-1]:
            want = int(np.searchsorted(e, v, side="right")) - 1
        assert po.bin_index(e, float(v)) == want, v


# ---------------------------------------------------------------------------
# scoring: brute force on tiny lattices
# ---------------------------------------------------------------------------
def brute_force(x, y, zp, e, g, sx, sy, sz, margin):
    """Membership by testing every sphere centre of the lattice; returns
    (pairs list of (key, y) in point order, ambiguous-mask)."""
    N, D, r, H = g.n_linear, g.diameter, g.radius, g.half_length
    centres = g.gso + D * np.arange(N)
    out, amb = [], np.zeros(len(x), bool)
    for i in range(len(x)):
        p = np.array([x[i] + sx, y[i] + sy, zp[i] + sz])
        if np.any(np.abs(np.abs(p) - H) < margin):
            amb[i] = True
            continue
        if not np.all(np.abs(p) < H):
            continue
        idx = [int(np.argmin(np.abs(centres - c))) for c in p]
        d = p - centres[idx]
        if abs(np.dot(d, d) - r * r) < margin * D or np.any(np.abs(np.abs(d) - D / 2) < margin):
            amb[i] = True
            continue
        if np.dot(d, d) <= r * r:
            out.append((idx[0] + idx[1] * N + idx[2] * N * N, e[i] / ((2.0 / 3.0) * D)))
    return out, amb


@pytest.mark.parametrize("H,D,seed", [(4.0, 1.0, 0), (5.0, 2.0, 1), (3.3, 0.7, 2), (8.0, 1.0, 3)])
def test_scoring_matches_brute_force(H, D, seed):
    g = po.geometry(H, D)
    rng = np.random.default_rng(seed)
    n = 4000
    x = rng.uniform(-1.3 * H, 1.3 * H, n)
    y = rng.uniform(-1.3 * H, 1.3 * H, n)
    z = rng.uniform(-2.5 * H, 0.3 * H, n)
    e = rng.uniform(1, 100, n)
    zp = z + g.half_length
    for _ in range(3):
        sx, sy, sz = rng.uniform(-g.half_length, g.half_length, 3)
        pk, py, keys, sums = po.score_oversample(x, y, zp, e, g, sx, sy, sz)
        want, amb = brute_force(x, y, zp, e, g, sx, sy, sz, 1e-9)
        assert not amb.any()
        assert [int(k) for k in pk] == [k for k, _ in want]
        np.testing.assert_array_equal(py, [v for _, v in want])
        # reduce: unique sorted keys, canonical sums close to exact sums
        assert list(keys) == sorted(set(int(k) for k in pk))
        for k, s in zip(keys, sums):
            assert s == pytest.approx(math.fsum(py[pk == k]), rel=1e-13)


def test_scoring_boundaries():
    g = po.geometry(4.0, 1.0)   # N = 8, H' = 4, gso = -3.5
    H = g.half_length
    # on the box faces: never in the box (strict <)
    x = np.array([H, -H, 0.5, 0.5, 0.5, 0.5])
    y = np.array([0.5, 0.5, H, -H, 0.5, 0.5])
    zp = np.array([0.5, 0.5, 0.5, 0.5, H, -H])
    pk, *_ = po.score_oversample(x, y, zp, np.ones(6), g, 0.0, 0.0, 0.0)
    assert len(pk) == 0
    # exactly on a sphere surface: dist == r^2 counts (<=)
    x = np.array([0.5 + 0.5, 0.5, 0.5])
    y = np.array([0.5, 0.5 - 0.5, 0.5])
    zp = np.array([0.5, 0.5, 0.5 + 0.25])
    pk, py, *_ = po.score_oversample(x, y, zp, np.ones(3), g, 0.0, 0.0, 0.0)
    # (1.0, .5, .5): a = 4.5 -> rint(4.5) = 4 (half to even), d = -0.5 -> in (on surface)
    assert len(pk) == 3
    assert int(pk[0]) == 4 + 4 * 8 + 4 * 64
    # half-way to even: a/D = 3.5 rounds to 4, 4.5 to 4, 5.5 to 6
    x = np.array([0.0, 1.0, 2.0])
    pk, *_ = po.score_oversample(x, np.full(3, 0.5), np.full(3, 0.5), np.ones(3), g, 0.0, 0.0, 0.0)
    assert [int(k) % 8 for k in pk] == [4, 4, 6]


def test_shift_table():
    g = po.geometry(2.5e3, 10.0)
    s = po.shift_table(1234, 0, 10000, True, True, g)
    assert np.all(s > -g.half_length) and np.all(s <= g.half_length)
    assert abs(s.mean()) < 0.02 * g.half_length
    s0 = po.shift_table(1234, 0, 50, False, False, g)
    assert np.all(s0 == 0.0)
    # distinct tracks and seeds give distinct streams
    assert not np.array_equal(po.shift_table(1234, 1, 5, True, True, g), s[:5])
    assert not np.array_equal(po.shift_table(1235, 0, 5, True, True, g), s[:5])


# ---------------------------------------------------------------------------
# Histogram known answers (tst/tests/HistogramTests.cu)
# ---------------------------------------------------------------------------
HIST_CASES = [
    # (type, nbins, lo, hi, keys, value, expected value binned, expected count)
    ("lin", 500, 0, 500, np.arange(100), 1.0, 1.0, 100),        # :36-84
    ("lin", 500, 0, 500, np.arange(100), 5.0, 5.0, 100),        # :86-134
    ("lin", 500, 0, 1000, np.ones(100), 5.0, 500.0, 1),         # :136-184
    ("lin", 30000, 1e-1, 1e7, np.ones(1000), 1000.0, 1e6, 1),   # :186-234
    ("log", 30000, 1e-1, 1e7, np.ones(1000), 1000.0, 1e6, 1),   # :386-434
    ("log", 500, 0, 500, np.arange(100), 1.0, None, 0),         # :236-284 (all edges NaN)
    ("log", 500, 0, 500, np.arange(100), 5.0, None, 0),         # :286-334
    ("log", 500, 0, 1000, np.ones(100), 5.0, None, 0),          # :336-384
]


@pytest.mark.parametrize("case", HIST_CASES)
def test_histogram_known_answers_oracle(case):
    kind, nb, lo, hi, keys, v, binned, count = case
    e = po.edges(kind, lo, hi, nb)
    k, s = po.reduce_pairs(keys.astype(np.uint64), np.full(len(keys), v))
    got = po.histogram_counts(e, s).astype(np.float64)
    np.testing.assert_array_equal(got, F.expected_bins(e, binned, count))


# ---------------------------------------------------------------------------
# integration known answers (tst/tests/SuperTrackVoxelConstrainedTests.cu)
# ---------------------------------------------------------------------------
def oracle_macro(name, n_tracks=None, n_os=None, seed=1234, shift=False):
    fx, (kind, nb, lo, hi), nt, nos, value, count = F.MACROS[name]
    nt = nt if n_tracks is None else n_tracks
    nos = nos if n_os is None else n_os
    x, y, z, e = fx()
    g = po.geometry(*F.geometry_of(name))
    ed = po.edges(kind, lo, hi, nb)
    X = np.tile(x, nt); Y = np.tile(y, nt); Z = np.tile(z, nt); E = np.tile(e, nt)
    off = np.arange(nt + 1, dtype=np.int64) * len(x)
    bins, _, _ = po.run(X, Y, Z, E, off, [g], None, nos, seed, shift, False, [ed])
    return ed, bins[0], value, count


@pytest.mark.parametrize("name", ["A1", "A2", "A4", "A6", "A8"])
def test_integration_known_answers_oracle(name):
    ed, bins, value, count = oracle_macro(name)
    np.testing.assert_array_equal(bins.astype(np.float64), F.expected_bins(ed, value, count))


def test_a3_plane_oracle():
    ed, bins, value, count = oracle_macro("A3")
    np.testing.assert_array_equal(bins.astype(np.float64), F.expected_bins(ed, value, count))


def test_a5_coincident_points_oracle():
    # A5 is 5 tracks x 1000 oversamples of the same unshifted track; every
    # oversample scores alike, so one track x 2 oversamples pins the answer.
    ed, bins, value, _ = oracle_macro("A5", n_tracks=1, n_os=2)
    np.testing.assert_array_equal(bins.astype(np.float64), F.expected_bins(ed, value, 3000 * 2))


def test_a8_keys_straddle_32_bits():
    x, y, z, e = F.a8()
    g = po.geometry(1.5e6, 10.0)
    pk, py, keys, sums = po.track_segments(x, y, z, e, g, 1, 0, 1, False, False)[0]
    assert [int(k) for k in pk] == F.A8_KEYS
    assert np.all(sums == 1.0000000000000002)


def test_a7_pi_estimate_oracle():
    x, y, z, e = F.a7()
    g = po.geometry(1.5e6, 1e3)
    ed = po.edges("lin", 1, 1e4, 10000)
    off = np.array([0, len(x)], np.int64)
    bins, _, _ = po.run(x, y, z, e, off, [g], None, 10000, 20240601, True, False, [ed])
    mid = (ed[:-1] + ed[1:]) / 2
    pi = 6 * float(np.sum(mid * bins[0])) / (20000 * 10000)
    assert 3.05 < pi < 3.3, pi


def test_energy_conservation():
    rng = np.random.default_rng(5)
    g = po.geometry(50.0, 10.0)
    n = 20000
    x = rng.normal(0, 15, n); y = rng.normal(0, 15, n); z = rng.uniform(-100, 0, n); e = rng.uniform(1, 50, n)
    for k, sh in enumerate(po.shift_table(9, 0, 4, True, True, g)):
        pk, py, keys, sums = po.score_oversample(x, y, z + g.half_length, e, g, *sh)
        assert math.fsum(sums) == pytest.approx(math.fsum(py), rel=1e-12)
        assert len(keys) == len(np.unique(pk))


def test_near_edge_matches_brute_force():
    # orc_near_edge against a direct scan of every edge (an edge within
    # rtol * |v| of v is adjacent to v for these rtol, since edges are
    # further apart than that), including values on and beside edges
    rng = np.random.default_rng(42)
    for kind, nb, lo, hi in [("lin", 3000, 0.01, 300.0), ("log", 1000, 1e-2, 1e3)]:
        e = po.edges(kind, lo, hi, nb)
        vals = np.concatenate([
            rng.uniform(0.005, 1.2 * hi, 4000),
            e[rng.integers(0, nb + 1, 500)],
            e[rng.integers(0, nb + 1, 500)] * (1 + rng.choice([-1, 1], 500) * rng.uniform(0, 3e-6, 500)),
            [np.nan, np.inf, -1.0, 0.0]])
        for rtol in (1e-6, 1e-12):
            for v in vals:
                want = bool(not np.isnan(v) and np.any(np.abs(v - e) <= rtol * abs(v)))
                assert po.near_edge(e, float(v), rtol) == want, (kind, v, rtol)


def test_run_counts_near_edge_sums():
    # the run driver's near-edge totals equal orc_near_edge applied to every
    # sum of every segment, and its bins equal the per-segment sums binned
    from supertrack_amd import workloads as W
    import supertrack_amd as st
    w = W.workload("cfg3")
    spec = dict(w["synth"], n_max=20000)
    x, y, z, e, off = st.synth_library(spec, 6, 0)
    g = po.geometry(2.5e3, 10.0)
    edges = [po.edges("lin", 0.01, 300.0, 3000), po.edges("log", 1e-2, 1e3, 1000)]
    bins, _, _, near = po.run(x, y, z, e, off, [g], None, 5, 7, True, True, edges, near_rtol=1e-3)
    sums = np.concatenate([s for t in range(6)
                           for _, _, _, s in po.track_segments(x[off[t]:off[t + 1]], y[off[t]:off[t + 1]],
                                                               z[off[t]:off[t + 1]], e[off[t]:off[t + 1]], g, 7, t,
                                                               5, True, True)])
    for h in range(2):
        np.testing.assert_array_equal(bins[h], po.histogram_counts(edges[h], sums))
        assert near[h] == sum(po.near_edge(edges[h], float(v), 1e-3) for v in sums)
    assert near[0] > 0 and near[1] > 0
