"""Build libbes_hip.so (HIP for gfx950) in-tree.

    python -m bes_amd.build

The library is compiled from bes_amd/csrc with hipcc for gfx950 only.  FP64
kernels are built with -ffp-contract=off so that their arithmetic is the
reference's, operation for operation; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbes_hip.so")
CLI = os.path.join(HERE, "bes_hip")
SOURCES = ["explicit.hip", "implicit.hip", "learned.hip", "amg.hip", "geometry.cpp", "output.cpp"]
HEADERS = ["context.hpp", "flux.hpp", "geometry.hpp", "kernels_common.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include"),
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libbes_hip.so")
    os.replace(LIB + ".tmp", LIB)
    # command-line driver over the C++ classes of include/bes/bes.hpp
    cli = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
           "-o", CLI + ".tmp", os.path.join(CSRC, "bes_main.cpp"), "-L" + HERE, "-lbes_hip",
           "-L/opt/rocm/lib", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cli), flush=True)
    res = subprocess.run(cli, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building the bes_hip driver failed")
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
