"""Builds libopenmoc_amd.so in-tree with hipcc for gfx950.

Usage: python -m openmoc_amd.build   (or openmoc_amd.build.build())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libopenmoc_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["tracks.cpp", "fluxfile.cpp", "segfile.cpp", "capi.cpp", "serve.cpp", "hip_solver.hip"]

# gfx950 only; packed FP32 instructions disabled (-packed-fp32-ops); hardware
# float atomics for the sweep tally (-munsafe-fp-atomics).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
HOST_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> Path:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
        else:
            cmd = [HIPCC, *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        _run(cmd)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
