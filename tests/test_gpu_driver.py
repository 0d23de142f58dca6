"""GPU tests of the C++ drivers over the UTC task shim.

The reference's golden frame file (benchapps/nbody/seq/nbody_outpos.txt) is the
output of `-n 1024 -l 100 -O file` with FP32 arithmetic; nbody_main must
reproduce it byte for byte with the reference arithmetic, and stay within the
reference's own GPU-vs-CPU envelope with the fast precisions.
"""
import hashlib
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "utc_amd" / "lib"
GOLDEN = Path(__file__).resolve().parent / "golden"
SEQ_SHA = "1484534f18e875397249e74b1a319161a7b0d70307ee4dffc1ea1683082ce1d3"

pytestmark = pytest.mark.gpu


def run(args, cwd):
    print("case:", " ".join(args), flush=True)
    sys.stdout.flush()
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=180)
    assert p.returncode == 0, p.stdout + p.stderr
    return p


def frames(path):
    lines = Path(path).read_text().splitlines()
    n = 1024
    f1 = np.array([[float(x) for x in l.split()] for l in lines[n + 2:2 * n + 2]])
    return lines, f1


@pytest.mark.parametrize("memtype", [0, 1, 2])
def test_nbody_main_reproduces_golden_file(tmp_path, memtype):
    run([str(BIN / "nbody_main"), "-f", "-n", "1024", "-l", "100", "-m", str(memtype), "-O", "out.txt", "-v"],
        tmp_path)
    data = (tmp_path / "out.txt").read_bytes()
    assert hashlib.sha256(data).hexdigest() == SEQ_SHA
    rec = (tmp_path / "time_record.txt").read_text().splitlines()
    assert rec[1] == "Average:" and rec[3] == "Max:" and rec[5] == "Min:"
    assert len(rec[0]) == 64  # four 16-wide fields


@pytest.mark.parametrize("precision", [1, 2, 3])
def test_nbody_main_fast_precisions_within_reference_envelope(tmp_path, precision):
    """Chaotic 100-step run: the reference's own CUDA output differs from its
    CPU output by up to 0.0591 (max) with median 1.9e-4; the fast paths must
    stay inside that envelope."""
    run([str(BIN / "nbody_main"), "-f", "-n", "1024", "-l", "100", "-P", str(precision), "-O", "o.txt"], tmp_path)
    lines, f1 = frames(tmp_path / "o.txt")
    _, g1 = frames(GOLDEN / "seq_nbody_outpos.txt")
    d = np.abs(f1 - g1)[:, :3]
    assert d.max() <= 0.0591
    assert np.median(d) <= 1e-3
    assert np.array_equal(f1[:, 3], g1[:, 3])


def test_nbody_main_double_frames(tmp_path):
    """FP64 build of the driver (the reference's FTYPE) with several frames."""
    run([str(BIN / "nbody_main"), "-n", "2000", "-b", "128", "-l", "20", "-o", "5", "-i", "1", "-O", "d.txt"],
        tmp_path)
    lines = (tmp_path / "d.txt").read_text().splitlines()
    n = 2048  # rounded up to a multiple of the block size
    assert len(lines) == 5 * (n + 1)
    # the reference labels frame k with k * timestep, not with the step count
    # (benchapps/nbody/utc/nbody_main.cc:236-238)
    assert lines[0] == "0.000000" and lines[n + 1] == "0.016000" and lines[4 * (n + 1)] == "0.064000"


def test_nbody_mtask_runs_concurrent_tasks(tmp_path):
    p = run([str(BIN / "nbody_mtask"), "-T", "4", "-n", "4096", "-l", "10", "-P", "2", "-v"], tmp_path)
    assert "Number of tasks: 4" in p.stdout
    rec = (tmp_path / "time_record.txt").read_text().splitlines()
    assert len(rec[0]) == 80  # five fields


def test_nbody_main_rejects_bad_options(tmp_path):
    p = subprocess.run([str(BIN / "nbody_main"), "-t", "2"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 1 and "require one thread" in p.stderr
