"""Build the HIP library libnuvort_hip.so in-tree for gfx950.

    python -m nuvort_amd.build          (or __graft_entry__.build())

Every source is compiled with hipcc; FP contraction is off so that FP64
expressions evaluate exactly as the reference's Fortran does, and packed
FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnuvort_hip.so")
SOURCES = ["k_util.hip", "k_velocity.hip", "k_topology.hip", "k_tree.hip", "k_walls.hip", "nv_api.cpp", "nv_static.cpp", "nv_io.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
# per source: the tree walks fetch the next node records at the top of each
# iteration, ahead of their use at its end; machine sinking would move those
# loads down to the use (to free their scalar registers) and expose their
# latency in every iteration
SOURCE_FLAGS = {"k_tree.hip": ["-mllvm", "-disable-machine-sink"]}


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        flags = FLAGS + SOURCE_FLAGS.get(src, [])
        cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + flags + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed on %s" % src)
        if verbose and r.stderr.strip():
            sys.stderr.write(r.stderr)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
