"""Build the engine's shared library (HIP for gfx950) in-tree.

    python -m blochlib_amd.build

Produces blochlib_amd/libblfid.so from csrc/blfid_kernels.hip and csrc/blfid_host.cpp.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libblfid.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 instructions stay disabled in every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, "blfid_kernels.hip"), os.path.join(CSRC, "blfid_host.cpp")]
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(HERE, "..", "include"), "-o", tmp,
           "-x", "hip", srcs[0], "-x", "hip", srcs[1]]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    print(OUT)
    sys.exit(0)
