"""The C ABI library loads and exports what include/supertrack_amd.h declares;
host-only entry points (geometry, synthetic tracks) behave.  CPU only: no
call here touches a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po
import supertrack_amd as st
from supertrack_amd import _capi, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "supertrack_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(st_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _capi.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_symbol_has_a_python_prototype():
    missing = [n for n in declared_functions() if n not in _capi.PROTOTYPES]
    assert not missing, missing


def test_version_and_error_string():
    lib = _capi.lib()
    assert b"gfx950" in lib.st_version()
    g = _capi.Geometry()
    rc = lib.st_geometry_make(C.c_double(1.0), C.c_double(-1.0), C.byref(g))
    assert rc == -1
    assert b"diameter" in lib.st_last_error()


@pytest.mark.parametrize("H,D", [(1.5e6, 1e3), (1.5e6, 10.0), (2.5e3, 2.0), (1e3, 1.0), (5e3, 100.0),
                                 (1.0, 0.3), (1.0, 5.0), (2.5e3, 3.7)])
def test_product_geometry_equals_oracle(H, D):
    a = st.SphericalGeometry(H, D)
    b = po.geometry(H, D)
    for f in ("half_length", "length", "diameter", "radius", "gso", "n_linear"):
        assert getattr(a.c, f) == getattr(b, f), f


def test_synth_is_deterministic_and_order_free():
    w = W.workload("cfg2")
    x, y, z, e, off, _ = W.make_library(w, n_tracks=6)
    x2, y2, z2, e2, off2, _ = W.make_library(w, n_tracks=6)
    for a, b in zip((x, y, z, e, off), (x2, y2, z2, e2, off2)):
        np.testing.assert_array_equal(a, b)
    # track 4 generated alone equals track 4 generated in the batch
    xs, ys, zs, es, o4 = st.synth_library(dict(w["synth"], entry_half_length=2500.0, voxel_half_length=2500.0), 1, 4)
    np.testing.assert_array_equal(xs, x[off[4]:off[5]])
    np.testing.assert_array_equal(es, e[off[4]:off[5]])


def test_synth_statistics_cfg1():
    w = W.workload("cfg1")
    x, y, z, e, off, _ = W.make_library(w)
    n = int(off[-1])
    assert 3500 < n < 6500          # ~700 events x (1 + 0.3 * 20)
    assert np.all((e >= 1.0) & (e <= 1000.0))
    assert 10 < np.median(e) < 22   # lognormal median 15 eV
    assert z.min() > -5000.0 - 200 and z.max() < -5000.0 + 10000.0 + 200
    # primary points share (x0, y0)
    prim = (x == x[0]) & (y == y[0])
    assert 500 < prim.sum() < 950


def test_synth_count_mode_exact_sizes():
    w = W.workload("cfg3")
    spec = dict(w["synth"])
    x, y, z, e, off = st.synth_library(spec, 5, 100)
    counts = np.diff(off)
    assert np.all(counts >= 1000) and np.all(counts <= 1000000)
    # the counts call and the fill agree track by track
    x1, _, _, _, off1 = st.synth_library(spec, 1, 102)
    assert len(x1) == counts[2]


def test_header_is_plain_c99(tmp_path):
    # the boundary is a C ABI: the header must compile as strict C
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    src = tmp_path / "h.c"
    src.write_text('#include "supertrack_amd.h"\nint main(void) { return 0; }\n')
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I",
                        os.path.join(ROOT, "include"), "-c", str(src), "-o", str(tmp_path / "h.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_library_chunks_reproduce_make_library():
    # the streamed library's runs, concatenated, are make_library's arrays
    w = W.workload("cfg5")
    w["chunk"] = 3
    w["synth"]["n_max"] = 20000
    nt, npts, chunks = W.library_chunks(w, 10, max_tracks=2)
    parts = list(chunks)
    x, y, z, e, off, gt = W.make_library(w, 10)
    assert nt == 10 and npts == int(off[-1])
    assert [len(p[5]) for p in parts] == [2, 1, 2, 1, 2, 1, 1]
    for k, ref in enumerate((x, y, z, e)):
        np.testing.assert_array_equal(np.concatenate([p[k] for p in parts]), ref)
    np.testing.assert_array_equal(np.concatenate([p[5] for p in parts]), gt)
    counts = np.concatenate([np.diff(p[4]) for p in parts])
    np.testing.assert_array_equal(counts, np.diff(off))
    # one geometry alone: its tracks, as a one-geometry library
    ng, npg, c1 = W.library_chunks(w, 10, only_geometry=1)
    p1 = list(c1)
    assert ng == int((gt == 1).sum()) and all((p[5] == 0).all() for p in p1)
    assert npg == int(sum(off[t + 1] - off[t] for t in range(10) if gt[t] == 1))
