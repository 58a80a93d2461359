"""Builds libspaghetty_amd.so (gfx950) in-tree with hipcc.

    python -m spaghetty_amd.build            # build
    python -m spaghetty_amd.build --asm      # also write the gfx950 assembly
                                             # of the kernels (for inspection)

Flags:
  --offload-arch=gfx950       CDNA4 only, no other targets
  -ffp-contract=off           the accumulate rounds x*f and y+(x*f) separately,
                              as the reference does (trunk/old_sort.f:535)
  -target-feature -packed-fp32-ops
                              no packed FP32 instructions anywhere
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libspaghetty_amd.so")

SOURCES = ["kernels.hip", "batch.hip", "serve.hip", "plan.cpp", "capi.cpp", "contract.cpp"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-result",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
    ]


def build(verbose: bool = False, asm: bool = False) -> str:
    """Compiles every source and links the shared library; returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"]
        cmd = [hipcc(), *common_flags(), *lang, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs,
           "-lrocblas"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if asm:
        for src in ("kernels.hip", "batch.hip"):
            out = os.path.join(objdir, os.path.splitext(src)[0] + f"-{ARCH}.s")
            cmd = [hipcc(), *common_flags(), "-x", "hip", "--cuda-device-only", "-S",
                   os.path.join(CSRC, src), "-o", out]
            subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, asm="--asm" in sys.argv))
