"""Build the native library graphpi_amd/libgraphpi_amd.so (host C++ + HIP for gfx950).

Host sources are compiled with g++ (OpenMP, no floating-point contraction so
the planner's cost model rounds like the reference's), HIP sources with hipcc
for gfx950, and everything is linked by hipcc into one shared library that
exports the C ABI of include/graphpi_amd.h and the C++ API of include/*.h.

    python -m graphpi_amd.build [-v]
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "graphpi_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgraphpi_amd.so")
INCLUDE = os.path.join(ROOT, "include")

HOST_SOURCES = ["host/pattern.cpp", "host/schedule.cpp", "host/dataloader.cpp", "host/gen.cpp",
                "host/graph.cpp", "host/capi.cpp", "host/kplan.cpp"]
HIP_SOURCES = ["hip/count.hip", "hip/special.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", "g++")
ARCH = os.environ.get("GP_OFFLOAD_ARCH", "gfx950")

CXXFLAGS = ["-std=c++17", "-O3", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
            "-I" + INCLUDE, "-I" + CSRC]
HIPFLAGS = ["-std=c++17", "-O3", "-fPIC", "--offload-arch=" + ARCH,
            # no packed fp32 instructions anywhere in device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + INCLUDE, "-I" + CSRC, "-Wno-unused-result"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    if verbose and r.stdout.strip():
        print(r.stdout)


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    steps = []
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        steps.append([CXX] + CXXFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        steps.append([HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), steps))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fopenmp", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-Wl,--no-undefined", "-Wl,-soname,libgraphpi_amd.so"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    """Compile the oracle (test infrastructure) and, where the reference tree is
    present, the reference build used to write golden fixtures."""
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", oracle, "liboracle.so"], verbose)
    if os.path.isdir("/root/reference/src") and shutil.which("g++"):
        _run(["make", "-s", "-j8", "-C", oracle, "ref"], verbose)


if __name__ == "__main__":
    v = "-v" in sys.argv
    print(build(verbose=v))
    build_oracle(verbose=v)
