"""Builds libgrns_hip.so in-tree with hipcc for gfx950.

FP64 code is compiled with contraction off (-ffp-contract=off): the bit patterns of the
interval bounds depend on every product and sum being rounded separately, as in the
reference's host code; the few fused multiply-adds the kernels need are explicit fma() calls.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgrns_hip.so")
WORKLOAD_LIB = os.path.join(HERE, "libgrns_workload.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["grns_lib.hip", "grns_batch.hip"]
HEADERS = ["grns_device.hpp", "grns_kernels.hpp", "grns_internal.hpp"]

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
]


def _compile(srcs, out, verbose):
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *[os.path.join(CSRC, s) for s in srcs]]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    """Compile libgrns_hip.so (the hot path) and libgrns_workload.so (synthetic inputs)."""
    _compile(SOURCES, LIB, verbose)
    _compile(["grns_workload.hip"], WORKLOAD_LIB, verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
