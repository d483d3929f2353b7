"""Build recipe of the HIP library (gfx950) and of the oracle.

    python -m itk_amd.build          # compiles libitk_mattes_hip.so in-tree
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libitk_mattes_hip.so")
SOURCES = ["mattes_kernels.hip", "mattes_fast.hip", "mattes_prep.hip", "mattes_capi.hip", "mattes_optim.hip",
           "mattes_sampling.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-pthread",
    # the FP64 coordinate path must round like the reference's x86-64 build
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_hip(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include")] + [
        os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def main(argv=None) -> int:
    build_hip(verbose=True)
    print("built", LIB)
    return 0


if __name__ == "__main__":
    sys.exit(main())
