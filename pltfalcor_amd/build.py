"""Builds the HIP library libplt_amd.so in-tree for gfx950 (and the CPU oracle used by tests).

Run as ``python -m pltfalcor_amd.build`` or through ``__graft_entry__.build()``. The library is
compiled from every ``csrc/*.hip`` file by one hipcc invocation; nothing is JIT-compiled at
import time, so the built ``.so`` is what ships to the GPU host.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libplt_amd.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    "-pthread",
    # Packed FP32 (v_pk_*_f32) stays off: hipcc would otherwise form it from scalar float code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libplt_amd.so)")


def build_library(verbose: bool = True) -> Path:
    sources = sorted(str(p) for p in (PKG / "csrc").glob("*.hip")) + sorted(str(p) for p in (PKG / "csrc").glob("*.cpp"))
    if not sources:
        raise RuntimeError("no HIP sources under pltfalcor_amd/csrc")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-o", str(tmp), *sources]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = True) -> Path:
    cmd = ["make", "-s", "-C", str(ROOT / "oracle")]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return ROOT / "oracle" / "liboracle.so"


def main() -> int:
    build_library()
    build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
