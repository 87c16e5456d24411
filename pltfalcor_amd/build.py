"""Builds the HIP library libplt_amd.so in-tree for gfx950 (and the CPU oracle used by tests).

Run as ``python -m pltfalcor_amd.build`` or through ``__graft_entry__.build()``. Every
``csrc/*.hip`` / ``csrc/*.cpp`` file is compiled to its own object (in parallel, under
``_build/``) and the objects are linked into one shared library; nothing is JIT-compiled at
import time, so the built ``.so`` is what ships to the GPU host.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libplt_amd.so"
OBJ_DIR = ROOT / "_build"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-pthread",
    # Packed FP32 (v_pk_*_f32) stays off: hipcc would otherwise form it from scalar float code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Multiply-adds fuse only within one source expression (the default, "fast", also fuses across
    # statements wherever the backend finds the pattern, which depends on the code around it): the
    # same device function then rounds the same way in every kernel it is inlined into, so the
    # solve megakernel and the split solve pass produce identical colours.
    "-ffp-contract=on",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libplt_amd.so)")


def _jobs() -> int:
    try:
        return max(1, min(16, int(os.environ.get("MAX_JOBS", "")) if os.environ.get("MAX_JOBS") else (os.cpu_count() or 4)))
    except ValueError:
        return 4


def build_library(verbose: bool = True) -> Path:
    sources = sorted((PKG / "csrc").glob("*.hip")) + sorted((PKG / "csrc").glob("*.cpp"))
    if not sources:
        raise RuntimeError("no HIP sources under pltfalcor_amd/csrc")
    OBJ_DIR.mkdir(exist_ok=True)
    hipcc = _hipcc()
    # the largest translation units first, so they start before the small ones
    sources.sort(key=lambda p: -p.stat().st_size)

    def compile_one(src: Path) -> Path:
        obj = OBJ_DIR / (src.name + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-c", "-o", str(obj), str(src)]
        if verbose:
            print("[build]", " ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs)]
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
