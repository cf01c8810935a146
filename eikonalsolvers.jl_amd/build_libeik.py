"""Compile libeik.so (all HIP sources under csrc/) for gfx950, in-tree next to the package."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "eikonalsolvers", "libeik.so")
SOURCES = ["eik_capi.hip", "eik_fim.hip", "eik_adj.hip", "eik_serve.hip", "eik_rays.hip", "eik_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP64 stencils follow the reference's operation order; no implicit FMA contraction
    "-ffp-contract=off",
    # atomics are issued as written: the kernels reduce across lanes themselves where that pays,
    # and the compiler's per-wave rewrite of an atomic waits for its return value on the spot
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    # the stencil combines comparisons with & and | on purpose (no short-circuit branches)
    "-Wno-bitwise-instead-of-logical",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC, *FLAGS, "-o", OUT + ".tmp", *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
