"""The batched library's host-side layout (supertrack_amd/csrc/layout.cpp),
compiled on its own with g++ and checked on the CPU by tests/layout_check.cpp:
z order and permutation per track, block starts, record offsets beyond 2^32
points, FP32 offsets within each block's error bound, fast-path eligibility,
and the sliding-window sphere bound against brute force."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build_layout_check(out_dir, extra=()):
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "layout_check")
    csrc = os.path.join(ROOT, "supertrack_amd", "csrc")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wall", "-pthread", *extra,
                    os.path.join(HERE, "layout_check.cpp"), os.path.join(csrc, "layout.cpp"),
                    os.path.join(csrc, "parallel.cpp"), "-o", exe], check=True)
    return exe


def test_layout_invariants():
    exe = build_layout_check(os.path.join(HERE, "_build"))
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ALL OK" in res.stdout
