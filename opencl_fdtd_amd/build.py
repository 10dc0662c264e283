"""Ahead-of-time build of libfdtd_hip.so for gfx950 (hipcc, in-tree).

python -m opencl_fdtd_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "fdtd_hip.hip")
OUT = os.path.join(HERE, "libfdtd_hip.so")
ARCH = os.environ.get("FDTD_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every product and sum rounded separately (same order as
# the reference templates).  Packed-FP32 instructions are disabled.
FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + FLAGS + ["-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
