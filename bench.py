#!/usr/bin/env python3
"""Benchmark of the VoxelConstrainedSphere superposition path on one MI355X.

Metric: deposits scored per second (one deposit = one point of a track in one
oversample), for the batched FP32 form, with the plain FP64 drop-in path and
the batched FP64 form measured beside it, and the FP64 CPU oracle timed on a
bounded sample of the same workload.

A step is one pass over the whole track library: every track x every
oversample, from the shifts to the histogram counts copied back to the host.
Inputs (the synthetic library) are resident in HBM before the timed region.

    python bench.py                      # default: cfg4 batched FP32 (+ cfg3 comparisons)
    python bench.py --workload cfg3 --steps 3 --warmup 1
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue peak of MI355X: 256 CUs x 4 SIMDs, one wave64 VALU instruction
# per SIMD every 4 cycles, 2.4 GHz peak engine clock.
VALU_PEAK_GINSTR = 256 * 4 * 2.4e9 / 4 / 1e9
# Kernel sources whose hash keys the committed counter summaries: counters
# measured on other sources are not used.
KERNEL_SOURCES = ["supertrack_amd/csrc/batch_kernel.hpp", "supertrack_amd/csrc/st_common.hpp",
                  "supertrack_amd/csrc/st_device.hpp", "supertrack_amd/csrc/st_internal.hpp",
                  "supertrack_amd/csrc/layout.hpp"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_source_hash():
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def load_counters(workload, n_tracks, n_os, params):
    """Per-launch counters of the batched kernel from the committed rocprofv3
    summary profiles/batch_kernel_<workload>_counters.txt ("key = value"
    lines), if it was measured on this workload, launch shape and kernel
    source.  Returns (dict, path) or (None, reason)."""
    path = os.path.join("profiles", f"batch_kernel_{workload}_counters.txt")
    full = os.path.join(ROOT, path)
    if not os.path.exists(full):
        return None, f"no {path}"
    c = {}
    with open(full) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#") or "=" not in line:
                continue
            k, v = (t.strip() for t in line.split("=", 1))
            c[k] = v
    want = {"workload": workload, "tracks": str(n_tracks), "oversamples": str(n_os),
            "table_slots": str(params["table_slots"]), "blocks_per_cu": str(params["blocks_per_cu"]),
            "points_per_thread": str(params["points_per_thread"]), "kernel_source_sha": kernel_source_hash()}
    for k, v in want.items():
        if c.get(k) != v:
            return None, f"{path} was measured with {k} = {c.get(k)}, this run has {v}"
    return c, path


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="cfg4")
    p.add_argument("--tracks", type=int, default=0, help="use only the first T tracks (0: all)")
    p.add_argument("--oversamples", type=int, default=0, help="override the workload's oversamples")
    p.add_argument("--no-compare", action="store_true", help="skip the plain-FP64 / batched-FP64 legs")
    p.add_argument("--plain-stride", type=int, default=16, help="plain FP64 leg: every k-th track of cfg3")
    p.add_argument("--cpu-stride", type=int, default=10, help="CPU oracle sample: every k-th track")
    p.add_argument("--cpu-oversamples", type=int, default=0, help="CPU sample oversamples (0: workload's)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--autotune", action="store_true")
    p.add_argument("--per-geometry", action="store_true",
                   help="mixed-geometry workloads: also run each geometry's tracks as its own library, "
                        "autotuned separately, and report the kernel times beside the one mixed launch")
    p.add_argument("--workers", type=int, default=4,
                   help="plain path: concurrent workers, one HIP stream each (the reference's [Run] Threads)")
    p.add_argument("--queued", action="store_true",
                   help="plain workloads: time the method with queued oversamples (one launch per track)")
    p.add_argument("--via-cli", action="store_true",
                   help="plain workloads: also time the C++ driver (supertrack CLI) on the same tracks")
    p.add_argument("--copy-probe", action="store_true",
                   help="only run a 1 GiB device-to-device copy 5 times (calibrates HBM byte counters)")
    return p.parse_args()


class Events:
    """HIP events recorded on a given stream (through the HIP runtime)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        ms = C.c_float(0)
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)


def build_library(w, n_tracks, precision):
    """Device library of the workload's first n_tracks tracks.  Mixed-geometry
    workloads are generated, laid out and uploaded one same-geometry run at a
    time (host arrays: None); the others return their host arrays too."""
    import supertrack_amd as st
    from supertrack_amd import workloads as W
    geoms = W.geometries(w)
    t0 = time.time()
    if len(geoms) > 1:
        nt, npts, chunks = W.library_chunks(w, n_tracks)
        log(f"library: {nt} tracks, {npts} points, streamed by geometry run")

        def logged(it):
            done = 0
            for c in it:
                yield c
                done += len(c[5])
                log(f"library: {done}/{nt} tracks appended ({time.time() - t0:.1f}s)")

        lib = st.TrackLibrary.streamed(nt, npts, logged(chunks), geoms, precision)
        log(f"library: {n_tracks} tracks, {npts} points; generate+layout+upload by geometry run "
            f"{time.time() - t0:.1f}s; {lib.info()}")
        return lib, None
    x, y, z, e, off, gt = W.make_library(w, n_tracks=n_tracks)
    t1 = time.time()
    lib = st.TrackLibrary(x, y, z, e, off, geoms, None, precision)
    t2 = time.time()
    log(f"library: {n_tracks} tracks, {int(off[-1])} points; generate {t1 - t0:.1f}s, "
        f"layout+upload {t2 - t1:.1f}s; {lib.info()}")
    return lib, (x, y, z, e, off, gt)


def time_batch(lib, w, n_os, steps, warmup, autotune, ev):
    """Returns (seconds per step, kernel ms per step, params, counts)."""
    import supertrack_amd as st
    from supertrack_amd import workloads as W
    hs = [st.Histogram(nb, lo, hi, kind) for kind, nb, lo, hi in w["hists"]]
    b = st.BatchRun(lib, hs, n_os, W.SEED, True, True, use_graph=True, autotune=autotune)
    stream = st.default_stream_handle()
    for _ in range(warmup):
        b.run()
        b.sync()
        for h in hs:
            h.fetch(reset=True)
    st.synchronize()
    kernel_ms, exact_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
        b.sync()
        counts = [h.fetch(reset=True) for h in hs]   # bins back on the host: part of a step
        k, x = b.times()                              # HIP events on the batch stream
        kernel_ms.append(k)
        exact_ms.append(x)
    st.synchronize()
    secs = (time.perf_counter() - t0) / steps
    params = b.params()
    params["exact_path_ms"] = float(np.mean(exact_ms))
    params.update(b.stats())
    b.free()
    for h in hs:
        h.free()
    return secs, float(np.mean(kernel_ms)), params, counts


def time_plain(w, host, stride, n_os, workers=1, queued=False):
    """Plain FP64 drop-in path: per track begin, n_os x (process, sort-reduce-bin), end.

    Tracks are split into contiguous ranges over `workers` workers (the
    reference's [Run] Threads); each worker has its own method, histogram,
    buffers and HIP stream, and the worker histograms are summed at the end
    (Histogram::ReduceVector)."""
    import threading
    import supertrack_amd as st
    from supertrack_amd import workloads as W
    x, y, z, e, off, gt = host
    tracks = list(range(0, len(off) - 1, stride))
    geoms = W.geometries(w)
    kind, nb, lo, hi = w["hists"][0]
    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    # tracks are loaded to the device outside the timed region (ingest is not part of the path)
    dev = []
    cap = 0
    for t in tracks:
        b, en = int(off[t]), int(off[t + 1])
        dev.append(st.DeviceArray.from_numpy(np.stack([x[b:en], y[b:en], z[b:en], e[b:en]])))
        cap = max(cap, en - b)
    n_w = max(1, min(workers, len(tracks)))
    bounds = [len(tracks) * k // n_w for k in range(n_w + 1)]
    state = []
    for k in range(n_w):
        stream = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(stream)) == 0
        h = st.Histogram(nb, lo, hi, kind)
        h.set_stream(stream)
        methods = [st.VoxelConstrainedSphereMethod(g, True, True) for g in geoms]
        for m in methods:
            m.set_stream(stream)
            if queued:
                m.set_queued(True)
        state.append(dict(stream=stream, h=h, methods=methods, trk=st.Track(capacity=cap),
                          pairs=st.VolumeEdepPair(cap), err=None))
    st.synchronize()

    process = st.lib().st_vcs_process_oversample
    hist_add = st.lib().st_hist_add

    def run_worker(k):
        S = state[k]
        try:
            trk, pairs, h = S["trk"], S["pairs"], S["h"]
            for j in range(bounds[k], bounds[k + 1]):
                t = tracks[j]
                n = int(off[t + 1] - off[t])
                base = dev[j].ptr.value
                for a, dst in enumerate((trk.c.x, trk.c.y, trk.c.z, trk.c.edep)):
                    assert hip.hipMemcpyAsync(C.c_void_p(dst), C.c_void_p(base + 8 * n * a), 8 * n, 3,
                                              S["stream"]) == 0
                trk.c.n = n
                m = S["methods"][int(gt[t]) if len(geoms) > 1 else 0]
                pairs.c.capacity = n
                m.allocate_track_process(trk, W.SEED, t, n_os)
                h.allocate_track_process(pairs)
                # the two per-oversample calls go straight to the C ABI (the
                # wrappers' Python overhead is of the order of a launch)
                tc, pc = C.byref(trk.c), C.byref(pairs.c)
                rc = 0
                for _ in range(n_os):
                    rc |= process(m.h, tc, pc)
                    rc |= hist_add(h.h, pc)
                if rc:
                    raise RuntimeError(f"plain path call failed: {st.lib().st_last_error()}")
                m.free_track_process()
                h.free_track_process()
        except Exception as ex:   # reported after the join
            S["err"] = ex

    t0 = time.perf_counter()
    threads = [threading.Thread(target=run_worker, args=(k,)) for k in range(n_w)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    st.synchronize()
    secs = time.perf_counter() - t0
    for S in state:
        if S["err"] is not None:
            raise S["err"]
    counts = sum(S["h"].counts() for S in state)
    deposits = sum(int(off[t + 1] - off[t]) for t in tracks) * n_os
    for S in state:
        S["pairs"].c.capacity = cap
        S["h"].free()
        for m in S["methods"]:
            m.free()
        hip.hipStreamDestroy(S["stream"])
    return deposits, secs, len(tracks), counts


def cpu_baseline(w, stride, n_os, rtol=None):
    """The FP64 oracle (OpenMP over (track, oversample) segments) on every
    stride-th track of the workload's library; sample track j uses shift
    counter j, as the batched run of the same sample does.  Returns (rate,
    threads, n_sample, deposits, seconds, sample) with sample = (host arrays,
    oracle bins, near-edge counts at rtol)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from supertrack_amd import workloads as W
    n_sample = (w["n_tracks"] + stride - 1) // stride
    x, y, z, e, off, gt = W.make_library(w, n_tracks=n_sample, t0=0, stride=stride)
    geoms = [po.geometry(h, d) for h, d in w["geometries"]]
    edges = [po.edges(kind, lo, hi, nb) for kind, nb, lo, hi in w["hists"]]
    deposits = int(off[-1]) * n_os
    mixed = len(geoms) > 1
    res = po.run(x, y, z, e, off, geoms, gt if mixed else None, n_os, W.SEED, True, True, edges, n_threads=0,
                 near_rtol=rtol)
    bins, secs, threads = res[:3]
    near = res[3] if rtol is not None else None
    return deposits / secs, threads, n_sample, deposits, secs, ((x, y, z, e, off, gt), bins, near)


def sample_parity(w, sample, n_os, rtol, stride, record_tracks=4, record_oversamples=4):
    """The CPU sample through the batched kernel: every bin against the
    oracle's (a count may differ only for a sum within rtol of an edge), and
    the largest per-sphere energy error on the recorded spheres of the first
    record_tracks sample tracks x record_oversamples oversamples."""
    import supertrack_amd as st
    import pyoracle as po
    from supertrack_amd import workloads as W
    (x, y, z, e, off, gt), want, near = sample
    mixed = len(w["geometries"]) > 1
    lib = st.TrackLibrary(x, y, z, e, off, W.geometries(w), gt if mixed else None, w["precision"])
    hs = [st.Histogram(nb, lo, hi, kind) for kind, nb, lo, hi in w["hists"]]
    b = st.BatchRun(lib, hs, n_os, W.SEED, True, True, use_graph=False)
    b.run()
    b.sync()
    bal = b.balance()
    got = [h.fetch() for h in hs]
    b.free()
    lib.free()
    for h in hs:
        h.free()
    diff = [int(np.abs(g.astype(np.int64) - v.astype(np.int64)).sum()) for g, v in zip(got, want)]
    # energy error on a recorded subset (keys must match exactly)
    k = min(record_tracks, len(off) - 1)
    sub = slice(0, int(off[k]))
    lib = st.TrackLibrary(x[sub], y[sub], z[sub], e[sub], off[:k + 1], W.geometries(w),
                          gt[:k] if mixed else None, w["precision"])
    hs = [st.Histogram(nb, lo, hi, kind) for kind, nb, lo, hi in w["hists"]]
    b = st.BatchRun(lib, hs, record_oversamples, W.SEED, True, True, use_graph=False)
    cap = int(off[k]) * record_oversamples + 1024
    b.record(cap)
    b.run()
    b.sync()
    seg, keys, sums = b.record_fetch(cap)
    b.free()
    lib.free()
    for h in hs:
        h.free()
    worst, n_sph, keys_ok = 0.0, 0, True
    geoms = [po.geometry(h, d) for h, d in w["geometries"]]
    for t in range(k):
        g = geoms[int(gt[t]) if mixed else 0]
        segs = po.track_segments(x[off[t]:off[t + 1]], y[off[t]:off[t + 1]], z[off[t]:off[t + 1]],
                                 e[off[t]:off[t + 1]], g, W.SEED, t, record_oversamples, True, True)
        for o, (_, _, wk, ws) in enumerate(segs):
            sel = seg == t * record_oversamples + o
            order = np.argsort(keys[sel], kind="stable")
            gk, gs = keys[sel][order], sums[sel][order]
            if not np.array_equal(gk, wk):
                keys_ok = False
                continue
            if len(ws):
                worst = max(worst, float((np.abs(gs - ws) / np.abs(ws)).max()))
            n_sph += len(ws)
    ok = keys_ok and all(d <= 2 * n for d, n in zip(diff, near)) and worst <= rtol and \
        bal["energy_in"] == bal["energy_out"] and bal["entries_created"] == bal["entries_retired"]
    return {"sample": f"every {stride}th track ({len(off) - 1} tracks) x {n_os} oversamples, all histograms",
            "bins_differ": diff, "near_edge": near,
            "rule": f"bins_differ <= 2 x near_edge (sums within {rtol:g} of an edge may change bin)",
            "max_rel_energy_error": worst, "tolerance": rtol,
            "energy_sample": f"first {k} sample tracks x {record_oversamples} oversamples, {n_sph} spheres, "
                             f"keys {'identical' if keys_ok else 'DIFFER'}",
            "conservation": bal, "ok": bool(ok)}


def copy_probe():
    """1 GiB device-to-device copies: a kernel of known traffic (1 GiB read, 1 GiB written)."""
    import supertrack_amd as st
    n = 1 << 30
    a, b = st.DeviceArray(n), st.DeviceArray(n)
    st.call("st_memset", a.ptr, C.c_int(1), C.c_size_t(n))
    for _ in range(5):
        t0 = time.perf_counter()
        st.call("st_memcpy_dtod", b.ptr, a.ptr, C.c_size_t(n))
        dt = time.perf_counter() - t0
        log(f"copy 1 GiB: {dt * 1e3:.3f} ms (host clock, synchronous)")
    a.free()
    b.free()


def plain_main(args, w, n_tracks, n_os):
    """Plain FP64 drop-in path (cfg1, cfg2): every track, one launch sequence per track."""
    import supertrack_amd as st
    from supertrack_amd import workloads as W
    host = W.make_library(w, n_tracks=n_tracks)
    total_points = int(host[4][-1])
    deposits = total_points * n_os
    for _ in range(args.warmup):
        time_plain(w, host, 1, n_os, args.workers, args.queued)
    times = []
    for _ in range(args.steps):
        d, secs, nt, counts = time_plain(w, host, 1, n_os, args.workers, args.queued)
        times.append(secs)
    secs = float(np.mean(times))
    value = deposits / secs
    mode = "queued oversamples (one launch per track)" if args.queued else "one launch sequence per oversample"
    log(f"{args.workload} plain, {mode}: {secs * 1e3:.1f} ms/step, {value:.3e} deposits/s")
    extra = {}
    if not args.queued:
        # the same tracks with queued oversamples, for comparison
        time_plain(w, host, 1, n_os, args.workers, True)
        _, qs, _, qcounts = time_plain(w, host, 1, n_os, args.workers, True)
        extra["queued_ms_per_step"] = qs * 1e3
        extra["queued_deposits_per_s"] = deposits / qs
        extra["queued_counts_equal"] = bool(np.array_equal(qcounts, counts))
        log(f"{args.workload} plain, queued oversamples: {qs * 1e3:.1f} ms, {deposits / qs:.3e} deposits/s")
    if args.via_cli:
        for queued in (False, True):
            ms = time_cli(w, host, n_os, args.workers, queued)
            tag = "cli_queued" if queued else "cli"
            extra.update({f"{tag}_run_ms": ms, f"{tag}_deposits_per_s": deposits / (ms * 1e-3)})
            log(f"{args.workload} plain via CLI{' (QueueOversamples)' if queued else ''}: {ms:.1f} ms, "
                f"{deposits / (ms * 1e-3):.3e} deposits/s")
        extra["cli_note"] = (f"supertrack CLI, {args.workers} workers: track file reads (prefetched), per-track "
                             "loops and the output file, in the program's own clock")
    cpu = None
    parity = None
    if not args.no_cpu:
        rate, threads, ns, cdep, csecs, sample = cpu_baseline(w, 1, n_os)
        cpu = {"value": rate, "unit": "deposits/s", "cores": threads, "kind": "port",
               "sample": f"{args.workload}: all {ns} tracks x {n_os} oversamples, {cdep:.3e} deposits in "
                         f"{csecs:.2f}s, FP64 oracle, OpenMP"}
        want = sample[1][0]
        diff = int(np.abs(counts.astype(np.int64) - want.astype(np.int64)).sum())
        parity = {"sample": f"all {ns} tracks x {n_os} oversamples (the timed workload)", "bins_differ": diff,
                  "rule": "bit-exact: bins_differ == 0", "ok": diff == 0}
    out = {
        "metric": "deposits scored/s on 1 MI355X, plain FP64 per-track path",
        "value": value, "unit": "deposits/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": secs * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n_tracks} tracks, {total_points} points, {n_os} oversamples, "
                               f"geometries {w['geometries']}, histograms {w['hists'][:1]}",
                   "deposits_per_step": deposits, "parallelism": f"1 GPU, {args.workers} workers (one stream each)",
                   "mode": mode},
        "roofline": {"bound": "hbm", "achieved": deposits * 32 / secs / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": deposits * 32 / secs / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "basis": "32 B per deposit / host-clock step time (per-track launch sequences)"},
        "cpu_baseline": cpu,
        "parity": parity,
    }
    out.update(extra)
    print(json.dumps(out))


def time_cli(w, host, n_os, workers, queued=False):
    """Runs the C++ driver (supertrack CLI) on the workload's tracks written
    as a track file; returns the run time it reports (ms)."""
    import subprocess
    import tempfile
    from supertrack_amd import trackfile
    x, y, z, e, off, gt = host
    (H, D), = w["geometries"]
    kind, nb, lo, hi = w["hists"][0]
    cli = os.path.join(ROOT, "supertrack_amd", "supertrack")
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "tracks"))
        os.makedirs(os.path.join(tmp, "out"))
        trackfile.write(os.path.join(tmp, "tracks", "tracks.sttrk"), x, y, z, e, off)
        macro = os.path.join(tmp, "bench.ini")
        with open(macro, "w") as f:
            f.write(f"""[Run]
Type = Folder
Path = {tmp}/tracks
FirstFile = -1
LastFile = -1
Threads = {workers}
Oversamples = {n_os}
RandomSeed = 1234

[Simulation]
Method = VoxelConstrainedSphere

[VoxelConstrainedSphere]
ScoringRegionHalfLength = {H}
ScoringSphereDiameter = {D}
ShiftTrack = true
ShiftTrackZ = true
QueueOversamples = {"true" if queued else "false"}

[Histogram]
Type = {kind}
NBins = {nb}
BinLower = {lo}
BinUpper = {hi}

[Output]
Path = {tmp}/out/
NamePrefix = bench_
""")
        res = subprocess.run([cli, macro], capture_output=True, text=True, timeout=600)
        if res.returncode != 0:
            raise RuntimeError(res.stdout + res.stderr)
        for line in res.stdout.splitlines():
            if line.startswith("Run time:"):
                return float(line.split()[2])
    raise RuntimeError("the CLI printed no run time")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    import supertrack_amd as st
    from supertrack_amd import workloads as W

    ev = Events()
    w = W.workload(args.workload)
    n_tracks = args.tracks or w["n_tracks"]
    n_os = args.oversamples or w["n_oversamples"]
    if w.get("path") == "plain":
        plain_main(args, w, n_tracks, n_os)
        return
    lib, host = build_library(w, n_tracks, w["precision"])
    info = lib.info()
    total_points = info["n_points"]
    deposits = total_points * n_os
    secs, kms, params, counts = time_batch(lib, w, n_os, args.steps, args.warmup, args.autotune, ev)
    value = deposits / secs
    # algorithmic bytes: every point once per oversample, plus the FP64 side
    # copy (32 B) of the points whose FP32 decision was rechecked
    bytes_per_dep = (16 + 32 * params["fp64_rechecks"] / deposits) if w["precision"] == st.ST_F32 else 32
    achieved = deposits * bytes_per_dep / (kms * 1e-3) / 1e9
    log(f"{args.workload} batched: {secs * 1e3:.1f} ms/step, kernel {kms:.1f} ms, {value:.3e} deposits/s, "
        f"params {params}")
    lib.free()
    per_geometry = None
    if args.per_geometry and len(w["geometries"]) > 1:
        per_geometry = {"mixed_kernel_ms": kms, "geometries": []}
        total_ms = 0.0
        for g, (hl, dia) in enumerate(w["geometries"]):
            ng, npg, chunks = W.library_chunks(w, n_tracks, only_geometry=g)
            lg = st.TrackLibrary.streamed(ng, npg, chunks, [W.geometries(w)[g]], w["precision"])
            _, kg, pg, _ = time_batch(lg, w, n_os, 1, 1, True, ev)
            lg.free()
            total_ms += kg
            per_geometry["geometries"].append({"half_length": hl, "diameter": dia, "tracks": ng, "points": npg,
                                               "kernel_ms": kg, "launch": {k: pg[k] for k in (
                                                   "table_slots", "blocks_per_cu", "points_per_thread")}})
            log(f"{args.workload} geometry {g} ({hl}, {dia}) alone, autotuned: {ng} tracks, kernel {kg:.1f} ms, "
                f"{pg['table_slots']} slots x {pg['blocks_per_cu']} wg/CU x {pg['points_per_thread']} pt/thread")
        per_geometry["sum_kernel_ms"] = total_ms
        log(f"{args.workload}: one mixed launch {kms:.1f} ms, three autotuned launches {total_ms:.1f} ms")
    # roofline: the kernel issues VALU instructions most of its time (it culls
    # most deposits without reading them and re-reads the rest from L2), so
    # the bound is VALU issue; HBM traffic and the algorithmic byte rate are
    # reported beside it.  Counters come from the committed rocprofv3 summary
    # of this same command, the time from this run's HIP events.
    counters, cref = load_counters(args.workload, n_tracks, n_os, params)
    roof = {"bound": "valu", "unit": "G VALU instr/s", "peak": VALU_PEAK_GINSTR, "achieved": None, "frac": None,
            "traffic": None, "counters": cref,
            "algorithmic_gbs": achieved,
            "algorithmic_basis": f"{bytes_per_dep:.6g} B per deposit x {deposits} deposits / kernel time: "
                                 "informational (most deposits are culled unread, the rest re-read from L2)"}
    if counters:
        valu = float(counters["SQ_INSTS_VALU"])
        traffic = (float(counters["FETCH_SIZE_KiB"]) * float(counters["fetch_scale"]) +
                   float(counters["WRITE_SIZE_KiB"]) * float(counters["write_scale"])) * 1024.0
        ach = valu / (kms * 1e-3) / 1e9
        roof.update({
            "achieved": ach, "frac": ach / VALU_PEAK_GINSTR, "traffic": traffic,
            "basis": f"SQ_INSTS_VALU {valu:.6g} per launch ({cref}) / {kms:.2f} ms kernel time (average of "
                     f"{args.steps} launches, HIP events on the batch stream); peak = 1024 SIMDs x 2.4 GHz / "
                     "4 cycles per wave64 VALU instruction",
            "valu_lane_instr_per_scored_point": valu * 64 / max(params["points_scored"], 1),
            "hbm": {"achieved_gbs": traffic / (kms * 1e-3) / 1e9, "peak_gbs": HBM_PEAK_GBS,
                    "frac": traffic / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "basis": "(FETCH_SIZE x fetch_scale + WRITE_SIZE x write_scale) per launch, scales from a "
                             "1 GiB copy"},
        })
    else:
        log(f"roofline counters not used: {cref}")

    extra = {}
    if not args.no_compare:
        w3 = W.workload("cfg3")
        n3 = args.tracks or w3["n_tracks"]
        if w3["synth"] == w["synth"] and w3["geometries"][0][0] == w["geometries"][0][0] and n3 == n_tracks:
            # cfg3 and cfg4 share the synthetic library: reuse the host arrays
            host3 = host
            lib3 = st.TrackLibrary(*host3[:5], W.geometries(w3), None, st.ST_F64)
        else:
            del host
            lib3, host3 = build_library(w3, n3, st.ST_F64)
        d3 = lib3.info()["n_points"] * w3["n_oversamples"]
        s64, k64, _, _ = time_batch(lib3, w3, w3["n_oversamples"], 1, 1, False, ev)
        lib3.free()
        lib3f = st.TrackLibrary(*host3[:5], W.geometries(w3), None, st.ST_F32)
        s32, k32, _, _ = time_batch(lib3f, w3, w3["n_oversamples"], 1, 1, False, ev)
        lib3f.free()
        pd, ps, pn, _ = time_plain(w3, host3, args.plain_stride, w3["n_oversamples"], args.workers)
        plain_rate = pd / ps
        extra = {
            "cfg3_batched_fp64_deposits_per_s": d3 / s64,
            "cfg3_batched_fp32_deposits_per_s": d3 / s32,
            "cfg3_plain_fp64_deposits_per_s": plain_rate,
            "cfg3_plain_fp64_sample": f"every {args.plain_stride}th track ({pn} tracks) x {w3['n_oversamples']} oversamples, "
                                      f"{args.workers} workers",
            "cfg3_speedup_batched_fp32_vs_plain_fp64": (d3 / s32) / plain_rate,
        }
        log(f"cfg3: batched fp64 {d3 / s64:.3e}/s, batched fp32 {d3 / s32:.3e}/s, plain fp64 {plain_rate:.3e}/s "
            f"({pd:.3e} deposits in {ps:.1f}s)")

    cpu = None
    parity = None
    if not args.no_cpu:
        cos = args.cpu_oversamples or n_os
        rtol = 1e-6 if w["precision"] == st.ST_F32 else 1e-12
        rate, threads, ns, cdep, csecs, sample = cpu_baseline(w, args.cpu_stride, cos, rtol)
        cpu = {"value": rate, "unit": "deposits/s", "cores": threads, "kind": "port",
               "sample": f"{args.workload}: every {args.cpu_stride}th track ({ns} tracks) x {cos} oversamples, "
                         f"{cdep:.3e} deposits in {csecs:.1f}s, FP64 oracle, OpenMP"}
        log(f"cpu oracle: {rate:.3e} deposits/s on {threads} threads")
        parity = sample_parity(w, sample, cos, rtol, args.cpu_stride)
        log(f"parity on the cpu sample: {parity}")
        del sample

    out = {
        "metric": "deposits scored/s on 1 MI355X, batched-FP32 vs plain-FP64; % of HBM peak",
        "value": value,
        "unit": "deposits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if w["precision"] == st.ST_F32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n_tracks} tracks, {total_points} points, {n_os} oversamples, "
                               f"geometries {w['geometries']}, histograms {w['hists']}",
                   "deposits_per_step": deposits, "parallelism": "1 GPU", "launch": params,
                   "fast_tracks": info["fast_tracks"], "exact_tracks": info["exact_tracks"],
                   "scored_fraction": params["points_scored"] / deposits,
                   "fp64_recheck_fraction_of_scored": params["fp64_rechecks"] / max(params["points_scored"], 1),
                   "in_sphere_fraction": params["points_in_sphere"] / deposits,
                   **({"per_geometry": per_geometry} if per_geometry else {})},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
