"""Build the HIP library libalbsim_hip.so for gfx950 with hipcc, in-tree.

Usage: python -m pyalbsim_amd.build   (also called by __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libalbsim_hip.so")
SOURCES = ["alb_common.hip", "alb_trace.hip", "alb_map.hip", "alb_gather.hip", "alb_adjoint.hip", "alb_radius.hip"]
ARCH = os.environ.get("ALB_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # products and sums round separately, in source order (matches NumPy)
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # native global FP64 atomic add for waveform bins
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["hipcc", *FLAGS, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
