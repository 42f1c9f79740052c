"""Builds the gfx950 HIP library libhipfoam.so in-tree (hipcc, no JIT cache).

Usage: python -m metal4foam_amd.build   (or __graft_entry__.build())
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhipfoam.so")
ARCH = os.environ.get("HIPFOAM_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 (v_pk_*_f32) stays off: the compiler would emit it for
    # float2-shaped code, and this build keeps to scalar VALU FP32.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # every float expression rounds as written (no contraction into FMA): the
    # results do not depend on how a kernel was specialised or inlined, and
    # the FP32 row sums round exactly like the CPU oracle's
    "-ffp-contract=off",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose: bool = False) -> str:
    out_tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *sources(), "-o", out_tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out_tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
