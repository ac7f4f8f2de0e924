"""End-to-end runs of the C++ interface layer (SuperTrackManager,
VoxelConstrainedSphereMethod, Histogram) through the `supertrack` command on
.ini macros shaped like the reference's tst/testMacros/testA*.ini."""
import os
import subprocess

import numpy as np
import pytest

import fixtures as F
import pyoracle as po
from supertrack_amd import trackfile

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "supertrack_amd", "supertrack")

MACRO = """; macro shaped like tst/testMacros/testA{name}.ini
[Run]
Type = Folder
Path = {folder}
FirstFile = -1 ; #-1 is the default case, will analyze whole folder
LastFile = -1 ;
Threads = {threads}
Oversamples = {nos}
RandomSeed = {seed} ; set to -1 for the current time

[Simulation]
Method = VoxelConstrainedSphere

[VoxelConstrainedSphere]
ScoringRegionHalfLength = {H}    ; side length in nanometers
ScoringSphereDiameter = {D}        ; sphere diameter in nanometers
ShiftTrack = {shift}
Batched = {batched}
Precision = {precision}

[Histogram]
Type = {kind}
NBins = {nb}
BinLower = {lo}
BinUpper = {hi}

[Output]
Path = {out}/
NamePrefix = test_{name}_
"""


def read_csv(path):
    rows = [l for l in open(path) if not l.startswith("#") and not l.startswith("low")]
    a = np.array([[float(v) for v in r.split(",")] for r in rows])
    meta = {}
    for l in open(path):
        if l.startswith("# ") and ":" in l:
            k, v = l[2:].split(":", 1)
            meta[k.strip()] = v.strip()
    return a[:, 0], a[:, 1], a[:, 2], meta


def run_macro(tmp_path, name, tracks, kind, nb, lo, hi, nos, H, D, shift=False, batched=False, seed=1234,
              threads=1, precision="f64", files=1):
    folder = tmp_path / "tracks"
    folder.mkdir(exist_ok=True)
    per = (len(tracks) + files - 1) // files
    for f in range(files):
        part = tracks[f * per:(f + 1) * per]
        if not part:
            continue
        xs, ys, zs, es = (np.concatenate([t[i] for t in part]) for i in range(4))
        off = np.concatenate([[0], np.cumsum([len(t[0]) for t in part])]).astype(np.int64)
        trackfile.write(str(folder / f"file{f:02d}.sttrk"), xs, ys, zs, es, off)
    out = tmp_path / "out"
    out.mkdir(exist_ok=True)
    macro = tmp_path / f"test{name}.ini"
    macro.write_text(MACRO.format(name=name, folder=folder, threads=threads, nos=nos, seed=seed, H=H, D=D,
                                  shift="true" if shift else "false", batched="true" if batched else "false",
                                  precision=precision, kind=kind, nb=nb, lo=lo, hi=hi, out=out))
    res = subprocess.run([CLI, str(macro)], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout + res.stderr
    return read_csv(str(out / f"test_{name}_{seed}.csv"))


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("name", ["A1", "A2", "A4", "A6", "A8"])
def test_reference_macros(tmp_path, name, batched):
    fx, (kind, nb, lo, hi), nt, nos, value, count = F.MACROS[name]
    H, D = F.geometry_of(name)
    trk = fx()
    lo_e, hi_e, counts, meta = run_macro(tmp_path, name, [trk] * nt, kind, nb, lo, hi, nos, H, D, batched=batched)
    edges = np.append(lo_e, hi_e[-1])
    np.testing.assert_array_equal(counts, F.expected_bins(edges, value, count))
    assert float(meta["Number of Tracks"]) == nt
    assert float(meta["Effective Number of Tracks"]) == nt * nos


def test_reference_seeds_and_workers_match_oracle(tmp_path):
    """Shifted tracks in several files and workers: each task's shift stream
    uses the seed str(RandomSeed) + str(worker) + str(task in worker)."""
    from supertrack_amd import workloads as W
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=7)
    tracks = [(x[off[t]:off[t + 1]], y[off[t]:off[t + 1]], z[off[t]:off[t + 1]], e[off[t]:off[t + 1]])
              for t in range(7)]
    lo_e, hi_e, counts, _ = run_macro(tmp_path, "S", tracks, "lin", 3000, 0.01, 300.0, 5, 2.5e3, 10.0, shift=True,
                                      seed=42, threads=3, files=2)
    # expected: worker split 3/2/2 over the 7 tasks in file order
    g = po.geometry(2.5e3, 10.0)
    ed = po.edges("lin", 0.01, 300.0, 3000)
    want = np.zeros(3000)
    split = [3, 2, 2]
    t = 0
    for worker, k in enumerate(split):
        for j in range(k):
            seed = int(f"42{worker}{j}")
            for (_, _, keys, sums) in po.track_segments(*tracks[t], g, seed, 0, 5, True, False):
                want += po.histogram_counts(ed, sums)
            t += 1
    np.testing.assert_array_equal(counts, want)
