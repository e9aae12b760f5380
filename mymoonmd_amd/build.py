"""Build libpmx.so (host set-up + HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python -m mymoonmd_amd.build   (or build() from Python)
The library is written next to this file so that it travels with the tree.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpmx.so")

SOURCES = [
    "host/mesh.cpp",
    "host/fe.cpp",
    "host/space.cpp",
    "capi.cpp",
    "solver.cpp",
    "batch.cpp",
    "nse.cpp",
    "pod.cpp",
    "io.cpp",
    "hip/kernels_la.hip",
    "hip/kernels_asm.hip",
    "hip/kernels_batch.hip",
    "hip/kernels_nse.hip",
    "hip/kernels_pod.hip",
]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-fopenmp",
    # no packed FP32 instructions (v_pk_*_f32) in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-parameter",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src):
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


CALLER_SRC = os.path.join(ROOT, "tests", "host", "pmx_caller.cpp")
CALLER = os.path.join(ROOT, "tests", "host", "pmx_caller")


def build_caller(verbose=False):
    """The C++ program of tests/host/pmx_caller.cpp, compiled with g++ against
    include/pmx.h and linked with libpmx.so (found next to this file through
    its rpath), as a program on the reference side would be."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROCM, "include"), CALLER_SRC, "-o", CALLER + ".tmp", "-L" + HERE, "-lpmx",
           "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../mymoonmd_amd:" + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("caller build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(CALLER + ".tmp", CALLER)
    if verbose:
        print("built", CALLER)
    return CALLER


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-fopenmp", "-Wl,--no-undefined",
           "-Wl,-rpath," + os.path.join(ROCM, "llvm", "lib"), "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_caller(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
