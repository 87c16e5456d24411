"""The library's host C++ (BVH builder, procedural scenes, environment-map importance, camera
jitter patterns, .spd reader) built for the host alone with ROCm's clang++ and
-fsanitize=address,undefined (clang, not this image's g++ 11, because the host code uses
_Float16), and run over ordinary, degenerate and malformed inputs (tests/sanitize/host_checks.cpp).
CPU only."""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "pltfalcor_amd" / "csrc"
CLANG = "/opt/rocm/llvm/bin/clang++"
SOURCES = ["bvh_build.cpp", "scene_procedural.cpp", "env_map.cpp", "sample_pattern.cpp", "spectral_tables.cpp"]


@pytest.mark.skipif(not Path(CLANG).exists(), reason="needs ROCm's clang++")
def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    exe = tmp_path / "host_checks"
    cmd = [CLANG, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-pthread", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{ROOT / 'include'}", f"-I{CSRC}", str(ROOT / "tests" / "sanitize" / "host_checks.cpp"),
           *[str(CSRC / s) for s in SOURCES], "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               TMPDIR=str(tmp_path))
    env.pop("LD_PRELOAD", None)
    run = subprocess.run([str(exe), str(ROOT / "tests" / "golden" / "tables")], capture_output=True, text=True,
                         timeout=600, env=env)
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert "host checks passed" in run.stdout


@pytest.mark.skipif(not Path(CLANG).exists(), reason="needs ROCm's clang++")
def test_sanitizer_build_reports_a_planted_overflow(tmp_path):
    """The same flags catch a heap overflow and a signed overflow (the runtime is linked and on)."""
    src = tmp_path / "planted.cpp"
    src.write_text("#include <cstdlib>\nint main(int argc, char**) {\n  int* p = new int[2];\n  int r = p[argc + 1];\n"
                   "  delete[] p;\n  int big = 2147483647;\n  return r + (big + argc) == 7;\n}\n")
    exe = tmp_path / "planted"
    build = subprocess.run([CLANG, "-O0", "-g", "-fsanitize=address,undefined", str(src), "-o", str(exe)],
                           capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ)
    env.pop("LD_PRELOAD", None)
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, env=env)
    assert run.returncode != 0 and "heap-buffer-overflow" in run.stderr
