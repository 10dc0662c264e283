"""Builds the launch-recording variants of the HIP library's host code
(test infrastructure): libfdtd_record.so and, with AddressSanitizer and
UBSan on the host code only, libfdtd_record_asan.so.  The kernels are
compiled as usual (their registration needs them) but never launched."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "opencl_fdtd_amd", "csrc", "fdtd_hip.hip")
PLAIN = os.path.join(HERE, "libfdtd_record.so")
ASAN = os.path.join(HERE, "libfdtd_record_asan.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
BASE = [HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-DFDTD_RECORD_LAUNCHES"]
# each -fsanitize= right after -Xarch_host: host code only
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libsan",
       "-fno-omit-frame-pointer"]


def asan_runtime() -> str:
    """Path of the clang AddressSanitizer runtime the ASan build links."""
    res = subprocess.run([os.path.join(os.path.dirname(HIPCC), "..", "llvm", "bin", "clang++"),
                          "-print-resource-dir"], capture_output=True, text=True, check=True).stdout.strip()
    return os.path.join(res, "lib", "linux", "libclang_rt.asan-x86_64.so")


def build(verbose: bool = False) -> None:
    for out, extra in ((PLAIN, []), (ASAN, SAN)):
        cmd = BASE + extra + ["-o", out + ".tmp", SRC]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
