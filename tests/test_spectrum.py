"""Spectrum post-processing of a finished histogram (CPUHistogramUtils,
utils.cc:80-117) through `supertrack --spectrum`: CPU only."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.environ.get("SUPERTRACK_AMD_CLI") or os.path.join(ROOT, "supertrack_amd", "supertrack")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="supertrack CLI not built")


def write_csv(path, edges, counts):
    with open(path, "w") as f:
        f.write("# Lineal energy histogram\n# Seed: 1\nlow_edge,high_edge,count\n")
        for lo, hi, c in zip(edges[:-1], edges[1:], counts):
            f.write(f"{lo:.17g},{hi:.17g},{c:.17g}\n")


def read_csv(path):
    rows = [l for l in open(path) if not l.startswith("#") and not l.startswith("low")]
    a = np.array([[float(v) for v in r.split(",")] for r in rows])
    return a[:, 0], a[:, 1], a[:, 2]


def low_edge(edges, i):
    """TAxis::GetBinLowEdge of a variable-bin axis (TH1 numbering)."""
    n = len(edges) - 1
    if 1 <= i <= n + 1:
        return edges[i - 1]
    return edges[0] + (i - 1) * ((edges[n] - edges[0]) / n)


def reference_steps(edges, counts, semilog):
    """The two loops as the reference runs them: TH1 bins 0..n-1 (underflow
    in, last bin untouched), single-precision normalisation."""
    n = len(edges) - 1
    c = np.concatenate([[0.0], counts, [0.0]])
    norm = np.float32(0.0)
    for i in range(n):
        v = c[i]
        c[i] = v / (low_edge(edges, i + 1) - low_edge(edges, i))
        norm = np.float32(float(norm) + v)
    for i in range(n):
        c[i] = c[i] / float(norm)
    if semilog:
        for i in range(n):
            c[i] = c[i] * ((low_edge(edges, i + 1) + low_edge(edges, i)) / 2)
    return c[1:n + 1]


@pytest.mark.parametrize("kind,nb,lo,hi,semilog", [("log", 1000, 1e-2, 1e3, False), ("log", 200, 1e-1, 1e2, True),
                                                   ("lin", 300, 0.01, 300.0, False), ("lin", 50, 1.0, 11.0, True)])
def test_spectrum_matches_reference_steps(tmp_path, kind, nb, lo, hi, semilog):
    rng = np.random.default_rng(nb)
    edges = po.edges(kind, lo, hi, nb)
    counts = rng.integers(0, 10**6, nb).astype(np.float64)
    src, dst = str(tmp_path / "h.csv"), str(tmp_path / "pdf.csv")
    write_csv(src, edges, counts)
    args = [CLI, "--spectrum", src, dst] + (["semilog"] if semilog else [])
    subprocess.run(args, check=True, capture_output=True, timeout=60)
    lo_e, hi_e, got = read_csv(dst)
    np.testing.assert_array_equal(lo_e, edges[:-1])
    np.testing.assert_array_equal(hi_e, edges[1:])
    want = reference_steps(edges, counts, semilog)
    np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)
    # the last bin is left as the reference leaves it (raw count)
    assert got[-1] == counts[-1]
    if not semilog:
        # what the normalisation fixes: density x width over bins 1..n-1 sums
        # to one, up to the single-precision sum of the counts
        widths = edges[1:] - edges[:-1]
        assert abs(np.sum(got[:-1] * widths[:-1]) - 1.0) < 1e-6


def test_spectrum_missing_input_fails(tmp_path):
    r = subprocess.run([CLI, "--spectrum", str(tmp_path / "nope.csv"), str(tmp_path / "o.csv")],
                       capture_output=True, timeout=60)
    assert r.returncode != 0
