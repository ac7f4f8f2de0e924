"""Runs the CPU suite under AddressSanitizer + UndefinedBehaviorSanitizer.

    python tests/sanitizers.py

1. the library layout checks (tests/layout_check.cpp + layout.cpp), g++;
2. the oracle (oracle/oracle.cpp), g++, with the oracle tests;
3. the whole host side of libsupertrack_amd.so and the supertrack driver
   (hipcc; device code is built as usual, only host code is instrumented),
   with the whole CPU suite (`pytest -m "not gpu"`).
Builds go to tests/_build/sanitize/.  Exits non-zero if any session reports
an error.  CPU only: nothing here runs on a GPU.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build", "sanitize")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
GCC_SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
# host-only instrumentation on hipcc lines: each flag behind -Xarch_host
HIP_HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


def run(cmd, env=None, **kw):
    print("+", " ".join(cmd), flush=True)
    return subprocess.run(cmd, env=env, **kw)


def gcc_runtime():
    return ":".join(subprocess.run(["gcc", f"-print-file-name={n}"], capture_output=True, text=True,
                                   check=True).stdout.strip() for n in ("libasan.so", "libubsan.so"))


def clang_runtime():
    c = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not c:
        raise RuntimeError("clang's ASan runtime is not installed")
    return c[-1]


def main():
    os.makedirs(OUT, exist_ok=True)
    results = {}
    env = dict(os.environ, **SAN_ENV)
    # 1. layout checks
    exe = os.path.join(OUT, "layout_check")
    csrc = os.path.join(ROOT, "supertrack_amd", "csrc")
    run(["g++", "-O1", "-std=c++17", "-pthread", *GCC_SAN, os.path.join(HERE, "layout_check.cpp"),
         os.path.join(csrc, "layout.cpp"), os.path.join(csrc, "parallel.cpp"), "-o", exe], check=True)
    results["layout"] = run([exe], env=env).returncode
    # 2. oracle
    orc = os.path.join(OUT, "liboracle.so")
    run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", *GCC_SAN,
         os.path.join(ROOT, "oracle", "oracle.cpp"), "-o", orc], check=True)
    e2 = dict(env, LD_PRELOAD=gcc_runtime(), SUPERTRACK_ORACLE_LIBRARY=orc)
    results["oracle"] = run([sys.executable, "-m", "pytest", "-q", "-x", "-s", "-m", "not gpu",
                             os.path.join(HERE, "test_oracle.py"),
                             os.path.join(HERE, "test_golden.py")], env=e2, cwd=ROOT).returncode
    # 3. library host code + driver, whole CPU suite
    sys.path.insert(0, ROOT)
    from supertrack_amd import build as B
    objs = []
    procs = []
    for src in B.sources():
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC, *B.COMMON, *B.DEVICE, *HIP_HOST_SAN, *lang, "-I", os.path.join(ROOT, "include"),
               "-I", B.HOST, "-c", src, "-o", obj]
        procs.append(subprocess.Popen(cmd))
    if any(p.wait() for p in procs):
        raise RuntimeError("sanitizer build of the library failed")
    lib = os.path.join(OUT, "libsupertrack_amd.so")
    run([HIPCC, f"--offload-arch={B.ARCH}", "-shared", "-fPIC", "-o", lib, *objs, "-lpthread"], check=True)
    # the driver's main is host-only C++: clang++ from the same toolchain
    cli = os.path.join(OUT, "supertrack")
    run(["/opt/rocm/llvm/bin/clang++", "-O1", "-std=c++17", *GCC_SAN, "-shared-libsan", "-I",
         os.path.join(ROOT, "include"), "-I", B.HOST, os.path.join(B.HOST, "supertrack_main.cc"), "-o", cli,
         "-L", OUT, "-lsupertrack_amd", f"-Wl,-rpath,{OUT}"], check=True)
    e3 = dict(env, LD_PRELOAD=clang_runtime(), SUPERTRACK_AMD_LIBRARY=lib, SUPERTRACK_AMD_CLI=cli)
    results["library"] = run([sys.executable, "-m", "pytest", "-q", "-x", "-s", "-m", "not gpu", HERE], env=e3,
                             cwd=ROOT).returncode
    print("sanitizer sessions:", results, flush=True)
    return 1 if any(results.values()) else 0


if __name__ == "__main__":
    sys.exit(main())
