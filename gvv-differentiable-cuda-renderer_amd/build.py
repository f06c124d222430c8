"""Builds libgvv_amd.so (the HIP renderer and its C ABI) for gfx950, in-tree.

Flags that the numeric contract depends on:
  -ffp-contract=off                        no FMA contraction: binary32 visibility is bit-exact
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division and square root
  -munsafe-fp-atomics                      hardware float atomics for the gradient scatter
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgvv_amd.so")
SOURCES = ["gvv_forward.hip", "gvv_backward.hip", "gvv_capi.hip", "gvv_conv.hip", "gvv_multi.hip", "gvv_calib.hip",
           "gvv_noalloc.hip", "gvv_tile.hip", "gvv_fit.hip"]
ARCH = os.environ.get("GVV_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *FLAGS, *[os.path.join(CSRC, s) for s in SOURCES], "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgvv_amd.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
