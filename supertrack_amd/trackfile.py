"""Reader/writer of the ROOT-free track library file (*.sttrk).

Layout (little endian): b"STTRACK1", uint64 n_tracks, uint64 n_points,
uint64 index[n_tracks] (cumulative end entry of each track, as the
reference's "Track index" tree holds), then float64 x, y, z, edep columns of
n_points each (nm, eV).  supertrack_amd/host/TrackFile.hh reads the same
format on the C++ side.
"""
from __future__ import annotations

import numpy as np

MAGIC = b"STTRACK1"


def write(path: str, x, y, z, edep, offsets) -> None:
    offsets = np.asarray(offsets, np.int64)
    if offsets[0] != 0:
        raise ValueError("offsets must start at 0")
    n = int(offsets[-1])
    cols = [np.ascontiguousarray(c, np.float64) for c in (x, y, z, edep)]
    if any(len(c) != n for c in cols):
        raise ValueError("column lengths differ from offsets[-1]")
    with open(path, "wb") as f:
        f.write(MAGIC)
        np.array([len(offsets) - 1, n], np.uint64).tofile(f)
        offsets[1:].astype(np.uint64).tofile(f)
        for c in cols:
            c.tofile(f)


def read(path: str):
    """Returns (x, y, z, edep, offsets)."""
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a track file")
        nt, n = (int(v) for v in np.fromfile(f, np.uint64, 2))
        index = np.fromfile(f, np.uint64, nt).astype(np.int64)
        cols = [np.fromfile(f, np.float64, n) for _ in range(4)]
    offsets = np.concatenate([[0], index]).astype(np.int64)
    return (*cols, offsets)
