"""Frozen vectors (tests/golden/*.npz, written by make_golden.py): the oracle
must keep reproducing them bit for bit (CPU), and so must the plain HIP path
(GPU)."""
import glob
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    d = np.load(path)   # plain arrays only; allow_pickle stays False
    return {k: d[k] for k in d.files}


def hist_spec(d):
    nb, lo, hi = d["hist"]
    return ("log" if int(d["hist_log"]) else "lin", int(nb), float(lo), float(hi))


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_golden(path):
    d = load(path)
    g = po.geometry(float(d["H"]), float(d["D"]))
    n_os = int(d["n_os"])
    segs = po.track_segments(d["x"], d["y"], d["z"], d["e"], g, int(d["seed"]), 0, n_os, bool(d["shift_xy"]),
                             bool(d["shift_z"]))
    kind, nb, lo, hi = hist_spec(d)
    ed = po.edges(kind, lo, hi, nb)
    counts = np.zeros(nb, np.uint64)
    for k, (pk, py, keys, sums) in enumerate(segs):
        np.testing.assert_array_equal(pk, d[f"pair_keys_{k}"])
        np.testing.assert_array_equal(py, d[f"pair_y_{k}"])
        np.testing.assert_array_equal(keys, d[f"keys_{k}"])
        np.testing.assert_array_equal(sums, d[f"sums_{k}"])
        counts += po.histogram_counts(ed, sums)
    np.testing.assert_array_equal(counts, d["counts"])
    assert d["counts"].sum() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_plain_path_reproduces_golden(path):
    import supertrack_amd as st
    d = load(path)
    g = st.SphericalGeometry(float(d["H"]), float(d["D"]))
    n_os = int(d["n_os"])
    m = st.VoxelConstrainedSphereMethod(g, bool(d["shift_xy"]), bool(d["shift_z"]))
    kind, nb, lo, hi = hist_spec(d)
    h = st.Histogram(nb, lo, hi, kind)
    trk = st.Track(d["x"], d["y"], d["z"], d["e"])
    pairs = st.VolumeEdepPair(len(d["x"]))
    m.allocate_track_process(trk, int(d["seed"]), 0, n_os)
    h.allocate_track_process(pairs)
    for k in range(n_os):
        m.process_track(trk, pairs)
        keys, ys = pairs.download()
        np.testing.assert_array_equal(keys, d[f"pair_keys_{k}"])
        np.testing.assert_array_equal(ys, d[f"pair_y_{k}"])
        h.sort_reduce_and_add_to_histogram(pairs)
    m.free_track_process()
    h.free_track_process()
    np.testing.assert_array_equal(h.counts(), d["counts"].astype(np.float64))
