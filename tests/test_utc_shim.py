"""CPU tests of the C++ UTC task shim: the self-test program, and the same
program rebuilt under ThreadSanitizer (host code only) to catch data races in
the worker-thread / job-queue machinery."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "utc_amd" / "lib"
UTC = ROOT / "utc_amd" / "utc"


def test_shim_self_test():
    exe = LIB / "test_utc_shim"
    assert exe.exists(), "build first: python -m utc_amd.build"
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "all checks passed" in p.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_shim_under_thread_sanitizer(tmp_path):
    exe = tmp_path / "test_utc_shim_tsan"
    src = [str(UTC / "test" / "test_utc_shim.cpp")] + [str(p) for p in sorted((UTC / "src").glob("*.cpp"))]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=thread", "-D__HIP_PLATFORM_AMD__",
           f"-I{UTC / 'include'}", "-I/opt/rocm/include", "-o", str(exe), *src,
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1")
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0, (p.stdout + p.stderr)[-4000:]
    assert "WARNING: ThreadSanitizer" not in p.stderr
