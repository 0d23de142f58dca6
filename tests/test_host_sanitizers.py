"""The generators' host code under AddressSanitizer + UndefinedBehaviorSanitizer
(CPU only): utc_amd/csrc/gen.cpp with tests/asan/gen_driver.cpp, built with g++
into a temporary directory and run."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_generators_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "gen_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           f"-I{ROOT / 'include'}", str(ROOT / "utc_amd" / "csrc" / "gen.cpp"),
           str(ROOT / "tests" / "asan" / "gen_driver.cpp"), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120,
                       env={"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout
