"""Known-answer fixtures of the reference's own test suite, re-expressed.

The reference's integration tests (tst/tests/SuperTrackVoxelConstrainedTests.cu)
run tracks written by tst/testTracks/tstTrackGenerator.cc through the whole
manager.  Their coordinates assume a kernel that moves z by -H'; the kernel
adds +H' (VoxelConstrainedSphereMethod.cu:152), so every fixture here is the
generator's track with z shifted by -2H' (so that z + H' lands where the
generator meant z - H' to land).  Expected counts are the tests' own.

All fixtures use H = 1.5e6 nm and D = 1e3 nm (N = 3000) except A8 (D = 10 nm,
N = 3e5).
"""
import numpy as np

H_REF = 1.5e6
SHIFT = -2 * H_REF


def _track(x, y, z, e):
    n = len(z)
    return (np.broadcast_to(np.asarray(x, np.float64), (n,)).copy(),
            np.broadcast_to(np.asarray(y, np.float64), (n,)).copy(),
            np.asarray(z, np.float64) + SHIFT,
            np.broadcast_to(np.asarray(e, np.float64), (n,)).copy())


def line_positions(start, stop, stride, inclusive=True):
    """The generator's `for (v = start; v <= stop; v += stride)` with double accumulation."""
    out = []
    v = float(start)
    while (v <= stop) if inclusive else (v < stop):
        out.append(v)
        v += stride
    return np.array(out, np.float64)


def a1():
    # tstTrackGenerator.cc:37-88: one point, edep = 2000/3 in integer arithmetic = 666 eV
    z = line_positions(0.5e3, 0.5e3 + 1e3, 1e3, inclusive=False)
    return _track(0.5e3, 0.5e3, z, float(2000 // 3))


def a2():
    # :90-141: 3000 points along z, one per sphere
    z = line_positions(0.5e3, 3e6 - 0.5e3, 1e3)
    return _track(0.5e3, 0.5e3, z, 2000.0 / 3)


def a3():
    # :143-195: a y-z plane of 3000 x 3000 points
    ys = line_positions(0.5e3 - 1.5e6, 3e6 - 0.5e3 - 1.5e6, 1e3)
    zs = line_positions(0.5e3, 3e6 - 0.5e3, 1e3)
    yy, zz = np.meshgrid(ys, zs, indexing="ij")
    return _track(0.5e3, yy.ravel(), zz.ravel(), 2000.0 / 3)


def a5_track():
    # :250-311: 1000 coincident points per sphere along z
    z = np.repeat(line_positions(0.5e3, 3e6 - 0.5e3, 1e3), 1000)
    return _track(0.5e3, 0.5e3, z, 2000.0 / 3)


def a6(step=50):
    """:313-405 intends points on the four x/y faces of the box; the
    generator's loops do not place them there (side 1 uses x = 500).  Here the
    intent: points with |x| = H' or |y| = H' exactly, over a grid of the other
    two coordinates (every `step`-th sphere centre), plus points on the z faces."""
    c = line_positions(0.5e3 - 1.5e6, 3e6 - 0.5e3 - 1.5e6, 1e3)[::step]
    zc = line_positions(0.5e3, 3e6 - 0.5e3, 1e3)[::step]
    xs, ys, zs = [], [], []
    for face in (-H_REF, H_REF):
        a, b = np.meshgrid(c, zc, indexing="ij")
        xs.append(np.full(a.size, face)); ys.append(a.ravel()); zs.append(b.ravel())
        xs.append(a.ravel()); ys.append(np.full(a.size, face)); zs.append(b.ravel())
    for zf in (0.0, 2 * H_REF):   # z faces in the generator's frame (z - H' = -H', +H')
        a, b = np.meshgrid(c, c, indexing="ij")
        xs.append(a.ravel()); ys.append(b.ravel()); zs.append(np.full(a.size, zf))
    x = np.concatenate(xs); y = np.concatenate(ys); z = np.concatenate(zs)
    return x, y, z + SHIFT, np.full(x.size, 2000.0 / 3)


def a7():
    # :407-461: 20000 points on the z axis at 10 nm pitch
    z = line_positions(0.0, 0.2e6, 10.0, inclusive=False)
    return _track(0.0, 0.0, z, 2000.0 / 3)


def a8():
    # :463-535: seven points whose keys straddle 2^31 and 2^32 (D = 10 nm)
    be = 1.5e6
    pts = [(5 - be, 5 - be, 5.0), (836465 - be, 71585 - be, 5.0), (836475 - be, 71585 - be, 5.0),
           (836485 - be, 71585 - be, 5.0), (1672955 - be, 143166 - be, 5.0),
           (1672965 - be, 143166 - be, 5.0), (5 - be, 5 - be, 1.5e6 - 5)]
    x = np.array([p[0] for p in pts]); y = np.array([p[1] for p in pts]); z = np.array([p[2] for p in pts])
    return x, y, z + SHIFT, np.full(7, 20.0 / 3)


A8_KEYS = [0, 2**31 - 2, 2**31 - 1, 2**31, 2**32 - 1, 2**32, 149999 * 300000 * 300000]

# (fixture, macro histogram (type, nbins, lo, hi), n_tracks, n_oversamples, expected value, expected count)
MACROS = {
    "A1": (a1, ("lin", 256, 1e-1, 1e2), 1, 1, 1.0, 1),
    "A2": (a2, ("lin", 256, 1e-1, 1e2), 1, 1, 1.0, 3000),
    "A3": (a3, ("lin", 256, 1e-1, 1e2), 1, 1, 1.0, 9000000),
    "A4": (a2, ("lin", 256, 1e-1, 1e2), 5, 1000, 1.0, 3000 * 1000 * 5),
    "A5": (a5_track, ("lin", 30000, 1e-1, 1e4), 5, 1000, 1000.0, 5 * 1000 * 3000),
    "A6": (a6, ("lin", 30000, 1e-1, 1e4), 1, 1, None, 0),
    "A8": (a8, ("lin", 10000, 0.01, 1e3), 1, 1, 1.0, 7),
}
GEOMETRY = {"A8": (1.5e6, 10.0)}


def geometry_of(name):
    return GEOMETRY.get(name, (1.5e6, 1e3))


def expected_bins(edges, value, count):
    """The tests' check: `count` in the bin with low <= value < high, 0 elsewhere."""
    out = np.zeros(len(edges) - 1, np.float64)
    if value is not None:
        for i in range(len(edges) - 1):
            if edges[i] <= value < edges[i + 1]:
                out[i] = count
    return out
