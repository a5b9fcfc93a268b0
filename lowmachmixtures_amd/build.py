"""Build recipe of the gfx950 library ``liblmm_hip.so`` (in-tree).

Compiles every ``csrc/*.hip`` translation unit for gfx950 with hipcc into one
shared library next to this file.  FP contraction is off so the FP64 kernels
keep the reference's rounding; packed FP32 instructions are disabled.
Run ``python -m lowmachmixtures_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblmm_hip.so")
ARCH = os.environ.get("LMM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def flags() -> list:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # Branches on workgroup-uniform values (component from blockIdx.y,
        # visc_type, colour) are emitted as plain scalar branches.  Without
        # this, a three-way uniform selection such as
        #   if (e == 0) ... else if (e == 1) ... else ...
        # is structurized so that the final case's value is never computed
        # on gfx950 (its VGPRs stay undefined and the load faults).
        "-mllvm", "-structurizecfg-skip-uniform-regions=true",
        "-Wno-unused-result",
        "-Wno-unused-value",
        "-I" + os.path.join(HERE, "..", "include"),
    ]


def build(verbose: bool = False) -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
