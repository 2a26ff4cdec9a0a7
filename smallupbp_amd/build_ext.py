"""Builds the HIP library libupbp_hip.so in-tree for gfx950 with hipcc.

All HIP code is compiled with FMA contraction off (the hit predicates must match
an IEEE FP32 evaluation bit for bit) and without packed-FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libupbp_hip.so")
SOURCES = ["capi.hip", "build.hip", "knn.hip", "lbvh.hip", "hashgrid.hip", "traverse.hip", "lighttrace.hip", "beamgrid.hip"]
ARCH = os.environ.get("UPBP_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def flags():
    return [
        "--offload-arch=%s" % ARCH,
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def build(verbose: bool = False) -> str:
    """Compile every source (always: this is the build check) and link the shared library."""
    from concurrent.futures import ThreadPoolExecutor

    cmds, objs = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(CSRC, s.replace(".hip", ".o"))
        cmds.append([hipcc()] + flags() + ["-c", src, "-o", obj])
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0) or os.cpu_count() or 1, 8))

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)

    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(compile_one, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
