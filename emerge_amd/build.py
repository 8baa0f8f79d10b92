"""Build libemerge_amd.so (HIP, gfx950) in-tree with hipcc.

``python -m emerge_amd.build`` or ``emerge_amd.build.build()``. The library lands next
to this file so it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libemerge_amd.so")
SOURCES = ["capi.hip", "ned2_assemble.hip", "ned2_ref.hip", "ned2_backward.hip", "csc_map.hip", "csc_freq.hip", "dof_select.hip", "robin.hip",
           "interp.hip", "heat.hip", "materials.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions stay disabled in every HIP build of this library.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-I", os.path.join(HERE, "..", "include")]


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *(extra or []), *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
