"""Builds libspca_mi355x.so (HIP kernels + host driver + C ABI) for gfx950 in-tree, and the
spca_console command-line front end linked against it.

Run as ``python 24am_amd/spca_mi355x/build.py``; __graft_entry__.build() calls build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(HERE, "libspca_mi355x.so")
CONSOLE = os.path.join(HERE, "spca_console")
SOURCES = ["spca_kernels.hip", "spca_host.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",  # keep the reference's operation order in elementwise code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-I" + os.path.join(REPO, "include"),
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    ctmp = CONSOLE + ".tmp"
    ccmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(REPO, "include"),
            os.path.join(CSRC, "spca_console.cpp"), "-o", ctmp, "-L" + HERE, "-l:libspca_mi355x.so",
            "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(ccmd), flush=True)
    subprocess.check_call(ccmd)
    os.replace(ctmp, CONSOLE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
