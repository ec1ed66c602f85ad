"""Builds libgdem.so (gfx950) in-tree with hipcc.

-ffp-contract=off keeps FP64 arithmetic free of fused multiply-adds so the
plain path reproduces the reference's -fmad=false results bit for bit;
packed FP32 instructions are disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgdem.so")
SOURCES = ["gdem_kernels.hip", "gdem_batch.hip", "gdem_noalloc.hip", "gdem_api.cpp", "gdem_batch.cpp", "gdem_mesh.cpp",
           "gdem_settings.cpp", "gdem_io.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


DEMO = os.path.join(HERE, "gdem_compat_demo")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    # the C++ caller of include/gdem_compat.h (host compiler only, links libgdem.so)
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "compat_demo.cpp"), "-o", DEMO + ".tmp", "-L", HERE, "-lgdem",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(DEMO + ".tmp", DEMO)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
