"""Build recipe for libsupertrack_amd.so (gfx950) and the oracle.

Compiles every HIP/C++ source under supertrack_amd/csrc with hipcc into one
shared library placed next to this file.  No CMake: the sources are few and
hipcc handles both device and host code.

    python -m supertrack_amd.build            # library + oracle
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "libsupertrack_amd.so")
CLI = os.path.join(HERE, "supertrack")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SUPERTRACK_ARCH", "gfx950")

# -ffp-contract=off: the FP64 path must round every product and sum the way
# the reference's source is written (no fused multiply-add).
# -packed-fp32-ops off: no v_pk_*_f32 instructions are generated.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]
HOST_ONLY = []


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    for name in sorted(os.listdir(HOST)):
        if name.endswith(".cc") and name != "supertrack_main.cc":
            out.append(os.path.join(HOST, name))
    return out


def _run(cmd, cwd=None):
    t0 = time.time()
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")
    return res.stdout, time.time() - t0


def build_library(jobs: int = 8, verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = ([HIPCC] + COMMON + DEVICE + lang +
               ["-I", os.path.join(ROOT, "include"), "-I", HOST, "-c", src, "-o", obj])
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            time.sleep(0.05)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out)
            sys.stderr.write(f"\nFAILED: {' '.join(cmd)}\n")
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("libsupertrack_amd build failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    # command line driver linked against the library (found next to it)
    tmp = CLI + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", HOST,
          os.path.join(HOST, "supertrack_main.cc"), "-o", tmp, "-L", HERE, "-lsupertrack_amd",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return LIB


def build_oracle() -> str:
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "all"])
    return os.path.join(odir, "liboracle.so")


def main():
    t0 = time.time()
    lib = build_library(verbose=True)
    orc = build_oracle()
    print(f"built {lib} and {orc} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
