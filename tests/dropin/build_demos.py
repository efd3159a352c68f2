"""Builds tests/dropin/dropin_demo.cpp (a program written against the GRNS API) with hipcc against
include/grns_dropin.hpp and grns_amd/libgrns_hip.so, once per moduli set: the reference's default
src/params.h and its shipped src/params/moduli-64-4096-bit/params.16m.256b.h (moduli from
tests/golden/params_sets.json).  Test infrastructure: called by __graft_entry__.build() and by the
CPU test that checks the drop-in header compiles; the binaries go to tests/dropin/_build/.

    python -m tests.dropin.build_demos
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SETS = {
    "params_h": "src/params.h",
    "16m_256b": "src/params/moduli-64-4096-bit/params.16m.256b.h",
}


def moduli_of(name: str) -> list[int]:
    with open(os.path.join(ROOT, "tests", "golden", "params_sets.json")) as f:
        sets = {e["file"]: e["moduli"] for e in json.load(f)}
    return sets[SETS[name]]


def binary(name: str) -> str:
    return os.path.join(OUT, name, "dropin_demo")


def build(verbose: bool = False, names=None) -> dict:
    """Writes each set's params.h and compiles the program; returns {name: binary path}."""
    sys.path.insert(0, ROOT)
    from grns_amd.moduli import params_header
    out = {}
    procs = []
    for name in names or SETS:
        d = os.path.join(OUT, name)
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "params.h"), "w") as f:
            f.write(params_header(moduli_of(name), note="GRNS " + SETS[name]))
        cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Wall", "-I", d, "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, "dropin_demo.cpp"), "-L", os.path.join(ROOT, "grns_amd"), "-lgrns_hip",
               "-Wl,-rpath,$ORIGIN/../../../../grns_amd", "-o", binary(name)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((name, subprocess.Popen(cmd)))
        out[name] = binary(name)
    failed = [n for n, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "dropin_demo " + ",".join(failed))
    return out


if __name__ == "__main__":
    print(build(verbose=True))
