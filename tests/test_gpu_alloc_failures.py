"""Allocation-failure paths of the C ABI, driven by st_debug_fail_allocation:
for every k up to the number of allocations a call makes, failing the k-th
one makes the call report ST_ENOMEM and release what it had allocated; the
objects made before keep working, and device memory returns to where it
started."""
import ctypes as C

import numpy as np
import pytest

import supertrack_amd as st
from supertrack_amd import _capi
from supertrack_amd import workloads as W

pytestmark = pytest.mark.gpu

ST_ENOMEM = -2


def _mem_free():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _each_failure(make):
    """Calls make() with the k-th allocation failing, k = 1, 2, ... until it
    succeeds; returns the number of allocations it makes."""
    k = 1
    while True:
        st.call("st_debug_fail_allocation", C.c_int64(k))
        try:
            make()
            ok = True
        except st.SuperTrackError as ex:
            assert ex.code == ST_ENOMEM, (k, str(ex))
            ok = False
        finally:
            st.call("st_debug_fail_allocation", C.c_int64(0))
        if ok:
            return k - 1
        k += 1
        assert k < 64


def test_every_allocation_failure_is_reported_and_released():
    w = W.workload("cfg3")
    x, y, z, e, off = st.synth_library(dict(w["synth"], n_max=30000), 6, 0)
    geoms = [st.SphericalGeometry(2.5e3, 10.0)]
    st.synchronize()
    free0 = _mem_free()

    def pairs():
        st.VolumeEdepPair(1000).free()

    def hist():
        st.Histogram(3000, 0.01, 300.0, "lin").free()

    def method():
        st.VoxelConstrainedSphereMethod(geoms[0], True, True).free()

    def track():
        st.Track(capacity=5000).free()

    def library():
        st.TrackLibrary(x, y, z, e, off, geoms, None, st.ST_F32).free()

    lib = st.TrackLibrary(x, y, z, e, off, geoms, None, st.ST_F32)
    h = st.Histogram(1000, 1e-2, 1e3, "log")

    def batch():
        b = st.BatchRun(lib, [h], 4, 5, True, True, use_graph=False)
        b.run()
        b.sync()
        b.free()

    big = int(off[1] - off[0])
    trk = st.Track(x[:big], y[:big], z[:big], e[:big])
    m = st.VoxelConstrainedSphereMethod(geoms[0], True, True)

    def begin_track():
        m.allocate_track_process(trk, 1, 0, 3)
        m.free_track_process()

    def sort_buffers():
        # pairs capacity above the one-workgroup sort: the radix-sort buffers
        made = []
        try:
            made.append(st.Histogram(100, 0.1, 10.0, "lin"))
            made.append(st.VolumeEdepPair(20000))
            made[0].allocate_track_process(made[1])
        finally:
            for o in made:
                o.free()

    counts = {}
    for name, fn in [("pairs", pairs), ("hist", hist), ("method", method), ("track", track),
                     ("library", library), ("batch", batch), ("begin_track", begin_track),
                     ("sort_buffers", sort_buffers)]:
        print(f"allocation failures: {name}", flush=True)
        counts[name] = _each_failure(fn)
    print(counts, flush=True)
    assert counts["pairs"] == 4 and counts["hist"] == 2 and counts["track"] == 4
    assert counts["library"] >= 8 and counts["batch"] >= 1 and counts["sort_buffers"] >= 7
    # what was made before the failures still works: a pass and its bins
    b = st.BatchRun(lib, [h], 4, 5, True, True)
    b.run()
    b.sync()
    assert h.fetch().sum() > 0
    b.free()
    for o in (m, trk, h, lib):
        o.free()
    st.synchronize()
    # nothing leaked (allowing for allocator granularity)
    assert abs(_mem_free() - free0) < (64 << 20), (free0, _mem_free())
