"""Build the gfx950 library (experiments_amd/libmi_rb.so) and the CPU oracle.

Every HIP source is compiled for gfx950 with hipcc; the library has a plain
C ABI (include/mi_rb.h) and does not depend on PyTorch.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmi_rb.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: no fused multiply-add, so the kernels round exactly like
# a serial evaluation of the same formulas.  The packed-fp32 target feature is
# disabled for every HIP compile.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]
HIP_SOURCES = ["k_sort.hip", "k_broad.hip", "k_narrow.hip", "k_solver.hip", "k_mlp.hip",
               "world.cpp"]
HOST_SOURCES = ["hull.cpp", "bullet_file.cpp", "obj_file.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src in HOST_SOURCES:
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-c", path, "-o", obj,
               "-I" + os.path.join(ROOT, "include")]
    else:
        cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_lib(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    sources = [s for s in HIP_SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


REFERENCE = "/root/reference"


def build_oracle(verbose=False):
    """The CPU oracle (test infrastructure); where the reference checkout is
    present, also its own CPU clip emulation into oracle/_ref (the checker's
    cross-check, never loaded by the product)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir(os.path.join(REFERENCE, "opencl")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "REF=" + REFERENCE])
    if verbose:
        print("built oracle")


def build(verbose=False):
    build_oracle(verbose)
    return build_lib(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
