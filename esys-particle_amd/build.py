"""Build libesp.so (HIP kernels + host orchestration + C-ABI) for gfx950, in place.

    python esys-particle_amd/build.py

-ffp-contract=off keeps every FP64 expression unfused (the laws round like the
reference's x86-64 build); packed FP32 instructions are disabled for all HIP code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libesp.so")
SOURCES = ["esp_kernels.hip", "esp_engine.hip", "esp_diff.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


GEOM_OUT = os.path.join(HERE, "libesp_geometry.so")


def build_geometry(verbose=False):
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", os.path.join(CSRC, "geometry.cpp"),
           "-o", GEOM_OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(GEOM_OUT + ".tmp", GEOM_OUT)


def build(verbose=False):
    build_geometry(verbose)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
