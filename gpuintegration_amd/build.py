"""Build the HIP library in-tree: gpuintegration_amd/libpagani.so (gfx950).

Also builds the test helper library tests/cxx/libreftests.so (the
reference's test integrands through the C++ header API) and the C interop
test program.  Every call recompiles.  Objects go to gpuintegration_amd/_build.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpagani.so")
TESTS_CXX = os.path.join(ROOT, "tests", "cxx")
REFTESTS_LIB = os.path.join(TESTS_CXX, "libreftests.so")
C_INTEROP_EXE = os.path.join(TESTS_CXX, "test_integration_from_c")

ARCH = os.environ.get("PAGANI_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 VALU forms are disabled for all device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I" + INCLUDE, "-I" + CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
]

# source -> extra flags
HIP_SOURCES = {
    "phases.hip": ["-ffp-contract=off"],
    "driver.hip": [],
    "genz_eval.hip": [],
    "interop.hip": [],
    "batch.hip": [],
    "checkpoint.hip": [],
    "mcubes.hip": [],
    "noalloc.hip": [],
}
CXX_SOURCES = {"rule.cpp": ["-O2", "-ffp-contract=off"],
               "mcubes_grid.cpp": ["-O2", "-ffp-contract=off"]}


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    return out


def _compile_hip(src, extra):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def _compile_cxx(src, extra):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run(["g++", "-std=c++17", "-fPIC", "-Wall", "-I" + INCLUDE, "-I" + CSRC, *extra, "-c",
          os.path.join(CSRC, src), "-o", obj])
    return obj


def build_library(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile_hip, s, f) for s, f in HIP_SOURCES.items()]
        futs += [ex.submit(_compile_cxx, s, f) for s, f in CXX_SOURCES.items()]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


HEADER_BENCH_LIB = os.path.join(HERE, "libpagani_header_bench.so")


def build_header_bench():
    """The benchmark's generic leg: a plain functor through the C++ header,
    compiled as a user translation unit against libpagani.so."""
    src = os.path.join(CSRC, "header_bench.hip")
    tmp = HEADER_BENCH_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", src, "-L" + HERE, "-lpagani",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, HEADER_BENCH_LIB)
    return HEADER_BENCH_LIB


def build_tests():
    src = os.path.join(TESTS_CXX, "reftests.hip")
    if not os.path.exists(src):
        return None
    tmp = REFTESTS_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", src, "-L" + HERE, "-lpagani",
          "-Wl,-rpath,$ORIGIN/../../gpuintegration_amd", "-o", tmp])
    os.replace(tmp, REFTESTS_LIB)
    csrc = os.path.join(TESTS_CXX, "test_integration_from_c.c")
    if os.path.exists(csrc):
        _run(["gcc", "-O2", "-Wall", "-I" + INCLUDE, csrc, "-L" + HERE, "-lpagani",
              "-Wl,-rpath," + "$ORIGIN/../../gpuintegration_amd", "-o", C_INTEROP_EXE, "-lm"])
    return REFTESTS_LIB


def build_all():
    build_library()
    with concurrent.futures.ThreadPoolExecutor(max_workers=2) as ex:
        futs = [ex.submit(build_header_bench), ex.submit(build_tests)]
        for f in futs:
            f.result()


if __name__ == "__main__":
    build_all()
    print(LIB)
    sys.exit(0)
