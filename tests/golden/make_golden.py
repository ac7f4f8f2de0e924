"""Writes tests/golden/*.npz: small tracks and the oracle's outputs for them.

The inputs are the first 400 points of synthetic tracks of the cfg1 and
cfg4 workloads plus the reference fixture A8; the outputs are the oracle's
per-oversample pair lists, unique keys, canonical sums and bin counts.
Regenerate only when the oracle changes on purpose:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import fixtures as F  # noqa: E402
import pyoracle as po  # noqa: E402
from supertrack_amd import workloads as W  # noqa: E402

CASES = {
    # name: (workload or fixture, geometry (H, D), n_os, seed, shift_xy, shift_z, histogram)
    "cfg1_head": ("cfg1", (2.5e3, 10.0), 4, 1234, True, True, ("lin", 3000, 0.01, 300.0)),
    "cfg4_head": ("cfg4", (2.5e3, 2.0), 24, 77, True, True, ("log", 1000, 1e-2, 1e3)),
    "a8": ("A8", (1.5e6, 10.0), 1, 1, False, False, ("lin", 10000, 0.01, 1e3)),
}


def inputs(src):
    if src == "A8":
        return F.a8()
    w = W.workload(src)
    spec = dict(w["synth"], entry_half_length=2500.0, voxel_half_length=2500.0)
    from supertrack_amd import synth_library
    x, y, z, e, off = synth_library(spec, 1, 3)
    # the first 400 points: the primary enters at the voxel's -z face
    return x[:400], y[:400], z[:400], e[:400]


def main():
    for name, (src, (H, D), n_os, seed, sxy, sz, (kind, nb, lo, hi)) in CASES.items():
        x, y, z, e = inputs(src)
        g = po.geometry(H, D)
        segs = po.track_segments(x, y, z, e, g, seed, 0, n_os, sxy, sz)
        ed = po.edges(kind, lo, hi, nb)
        out = dict(x=x, y=y, z=z, e=e, H=H, D=D, n_os=n_os, seed=seed, shift_xy=int(sxy), shift_z=int(sz),
                   hist=np.array([nb, lo, hi]), hist_log=int(kind == "log"))
        counts = np.zeros(nb, np.uint64)
        for k, (pk, py, keys, sums) in enumerate(segs):
            out[f"pair_keys_{k}"] = pk
            out[f"pair_y_{k}"] = py
            out[f"keys_{k}"] = keys
            out[f"sums_{k}"] = sums
            counts += po.histogram_counts(ed, sums)
        out["counts"] = counts
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, len(x), "points", int(counts.sum()), "spheres")


if __name__ == "__main__":
    main()
