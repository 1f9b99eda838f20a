"""Compile the HIP engine for gfx950 into origami_amd/_lib/liborigami_gpu.so.

Run as ``python -m origami_amd.build``.  Cross-compiles without a GPU.
Packed FP32 instructions are disabled for every kernel
(``-target-feature -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "liborigami_gpu.so")

SOURCES = ["engine.cpp", "pair_kernels.hip", "md_kernels.hip", "sasa_kernel.hip", "sasa_terms.hip",
           "surrogate.hip", "lbfgs.hip"]

# Per-file additions.  The SASA JVP terms take one-hot tangents whose zero
# products the compiler may drop only under finite-math, no-signed-zero
# assumptions; for finite values the terms are unchanged (see the file).
EXTRA = {"sasa_terms.hip": ["-ffinite-math-only", "-fno-signed-zeros"]}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-Wno-unused-command-line-argument", "-Wno-unused-value", "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "-x", "hip", "-c", path, "-o", obj] + flags() + EXTRA.get(src, [])
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=gfx950", "-fPIC"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
