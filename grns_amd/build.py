"""Builds libgrns_hip.so in-tree with hipcc for gfx950.

FP64 code is compiled with contraction off (-ffp-contract=off): the bit patterns of the
interval bounds depend on every product and sum being rounded separately, as in the
reference's host code; the few fused multiply-adds the kernels need are explicit fma() calls.
"""
from __future__ import annotations

import os
import shutil
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


def _run(cmds, verbose):
    """Run the commands in parallel; raise if any fails."""
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
    failed = [cmd for cmd, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])


def _snapshot(obj_dir: str) -> str:
    """Copy of the sources (csrc/ and include/) that the compilers read: hipcc reads a source and
    its headers once for the device code and again, minutes later, for the host code, so a file
    edited in between would give a library whose host side launches kernels its device code lacks."""
    src = os.path.join(obj_dir, "src")
    if os.path.isdir(src):
        shutil.rmtree(src)
    shutil.copytree(CSRC, os.path.join(src, "grns_amd", "csrc"))
    shutil.copytree(os.path.join(os.path.dirname(HERE), "include"), os.path.join(src, "include"))
    return os.path.join(src, "grns_amd", "csrc")


def build(verbose: bool = False) -> str:
    """Compile libgrns_hip.so (the hot path) and libgrns_workload.so (synthetic inputs) from one
    snapshot of the sources: every source to an object in parallel, then one link per library."""
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    csrc = _snapshot(obj_dir)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    units = SOURCES + ["grns_workload.hip"]
    objs = {u: os.path.join(obj_dir, u.replace(".hip", ".o")) for u in units}
    _run([[HIPCC, *compile_flags, "-c", "-o", objs[u], os.path.join(csrc, u)] for u in units], verbose)
    links = [(LIB, [objs[u] for u in SOURCES]), (WORKLOAD_LIB, [objs["grns_workload.hip"]])]
    _run([[HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp", *o] for out, o in links], verbose)
    for out, _ in links:
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
