"""Build recipe of the HIP library (gfx950) -- in-tree, no JIT cache."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpoongback_hip.so")

RESOURCES = os.path.join(HERE, "kernel_resources.txt")  # the compiler's per-kernel register/scratch report
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # bit-reproducible banded solves: no FP contraction anywhere
    "-ffp-contract=off",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = ["capi.hip", "bspline.cpp"]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-Rpass-analysis=kernel-resource-usage", *SOURCES, "-o", OUT]
    res = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed for libpoongback_hip.so")
    with open(RESOURCES, "w") as f:
        f.write("\n".join(l for l in res.stderr.splitlines() if "kernel-resource-usage" in l) + "\n")
    if verbose:
        sys.stderr.write(" ".join(cmd) + "\n")
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
