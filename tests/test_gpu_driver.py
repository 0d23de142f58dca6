"""GPU tests of the C++ drivers over the UTC task shim (nbody_main,
nbody_mtask, nbody_mgpu).

The reference's golden frame file (benchapps/nbody/seq/nbody_outpos.txt) is the
output of `-n 1024 -l 100 -O file` with FP32 arithmetic; nbody_main must
reproduce it byte for byte with the reference arithmetic, and stay within the
reference's own GPU-vs-CPU envelope with the fast precisions.
"""
import ctypes as C
import hashlib
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

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
    """Chaotic 100-step run: the reference's own CUDA output
    (cuda/nbody_outpos.txt) differs from its CPU output (the golden frame) by
    max 0.0591, 99th percentile 0.0112 and median 0.9e-4 over the 3072
    coordinates.  In a chaotic run equally valid FP32 roundings end far apart,
    and the maximum is the noisiest of the three: over the fast paths' tiles
    and both pair-scale forms (rsq alone or with a Newton step), 28 runs on
    MI355X deviated from the CPU frame by max 0.018-0.077, 99th percentile
    0.0055-0.0138 and median (0.9-1.5)e-4.  Each statistic must stay within
    twice the CUDA output's."""
    run([str(BIN / "nbody_main"), "-f", "-n", "1024", "-l", "100", "-P", str(precision), "-O", "o.txt"], tmp_path)
    lines, f1 = frames(tmp_path / "o.txt")
    _, g1 = frames(GOLDEN / "seq_nbody_outpos.txt")
    _, c1 = frames(GOLDEN / "cuda_nbody_outpos.txt")
    d = np.abs(f1 - g1)[:, :3]
    env = np.abs(c1 - g1)[:, :3]
    assert d.max() <= 2 * env.max()
    assert np.percentile(d, 99) <= 2 * np.percentile(env, 99)
    assert np.median(d) <= 2 * np.median(env)
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


PRESETS = json.loads((GOLDEN / "preset_frames_fp64.json").read_text())


@pytest.mark.parametrize("n", sorted({r["n"] for r in PRESETS["frames"]}))
def test_nbody_main_presets_reproduce_reference_driver(tmp_path, n):
    """The UTC driver (FP64, the reference's arithmetic on the GPU) for every
    preset -i 0..6 at N: byte for byte the frame file of the reference's own
    sequential driver (tests/golden/preset_frames_fp64.json)."""
    for rec in [r for r in PRESETS["frames"] if r["n"] == n]:
        out = tmp_path / f"p{rec['preset']}.txt"
        run([str(BIN / "nbody_main"), "-n", str(n), "-i", str(rec["preset"]), "-l", str(rec["steps"]),
             "-O", str(out)], tmp_path)
        assert hashlib.sha256(out.read_bytes()).hexdigest() == rec["sha256"], rec["preset"]


def _mtask_oracle_frames(ntasks, n, steps, dtype):
    """The oracle's runs of the multi-task driver's systems: task i starts from
    the i-th consecutive SHELL set of one rand() stream (as NbodyInit runs task
    after task, benchapps/nbody/utc/nbody_mtask.cc:185-199)."""
    prm = O.demo_params(0, n)
    O.lib().orc_srand(1)
    f = O.lib().orc_randomize_f32 if dtype == np.float32 else O.lib().orc_randomize_f64
    out = []
    for _ in range(ntasks):
        pos = np.zeros((n, 4), dtype)
        vel = np.zeros((n, 4), dtype)
        f(1, O._ptr(pos), O._ptr(vel), None, C.c_float(prm["clusterScale"]), C.c_float(prm["velocityScale"]),
          n, 1)
        eps2 = dtype(prm["softening"]) * dtype(prm["softening"])
        if dtype == np.float32:
            p, _ = O.run_f32(pos, vel, steps, eps2, dtype(prm["damping"]), dtype(prm["timestep"]), nthreads=16)
        else:
            p, _ = O.run_f64(pos, vel, steps, eps2, float(prm["damping"]), float(prm["timestep"]), nthreads=16)
        out.append((pos, p))
    return out, prm


def test_nbody_mtask_tasks_match_oracle(tmp_path):
    """-T 4 -n 4096 -l 10 in FP32: each task's frame file is byte for byte the
    oracle's run of that task's initial conditions (the reference arithmetic);
    with -P 2 (mixed) every task stays within 2e-5 of it (two steps of the
    frame file's %.5f)."""
    T, n, steps = 4, 4096, 10
    ref, prm = _mtask_oracle_frames(T, n, steps, np.float32)
    run([str(BIN / "nbody_mtask"), "-f", "-T", str(T), "-n", str(n), "-l", str(steps), "-O", "ref"], tmp_path)
    run([str(BIN / "nbody_mtask"), "-f", "-T", str(T), "-n", str(n), "-l", str(steps), "-P", "2", "-O", "mix"],
        tmp_path)
    for i, (p0, p1) in enumerate(ref):
        want = tmp_path / f"oracle.{i}"
        O.write_frames(want, np.stack([p0, p1]), prm["timestep"])
        assert (tmp_path / f"ref.{i}").read_bytes() == want.read_bytes(), i
        got = np.array([[float(x) for x in l.split()] for l in
                        (tmp_path / f"mix.{i}").read_text().splitlines()[n + 2:2 * n + 2]])
        assert np.abs(got - p1).max() <= 2e-5, i
    # the tasks' systems differ (successive rand() streams)
    assert not np.array_equal(ref[0][0], ref[1][0])


def test_nbody_mtask_runs_concurrent_tasks(tmp_path):
    p = run([str(BIN / "nbody_mtask"), "-T", "4", "-n", "4096", "-l", "10", "-P", "2", "-v"], tmp_path)
    assert "Number of tasks: 4" in p.stdout
    rec = (tmp_path / "time_record.txt").read_text().splitlines()
    assert len(rec[0]) == 80  # five fields


@pytest.mark.parametrize("threads,memtype", [(4, 0), (3, 1), (8, 2)])
def test_nbody_mgpu_reproduces_golden_file(tmp_path, threads, memtype):
    """The multi-GPU driver (BodySystemMGPU: each task thread advances its slice
    of the bodies against all of them and the slices meet in host buffers every
    step; here the threads share the one GPU) in FP32: byte for byte the
    reference's golden frame file, like the single-task driver."""
    run([str(BIN / "nbody_mgpu"), "-f", "-t", str(threads), "-m", str(memtype), "-n", "1024", "-l", "100",
         "-O", "out.txt", "-v"], tmp_path)
    data = (tmp_path / "out.txt").read_bytes()
    assert hashlib.sha256(data).hexdigest() == SEQ_SHA
    rec = (tmp_path / "time_record.txt").read_text().splitlines()
    assert len(rec[0]) == 64  # four 16-wide fields


def test_nbody_mgpu_presets_fp64(tmp_path):
    """FP64 (the reference driver's FTYPE), uneven slices (N = 4096 over 3 and
    over 7 threads): byte for byte the reference's sequential driver's frame
    files (tests/golden/preset_frames_fp64.json)."""
    for threads, preset in ((3, 2), (7, 5)):
        rec = next(r for r in PRESETS["frames"] if r["n"] == 4096 and r["preset"] == preset)
        out = tmp_path / f"m{threads}.txt"
        run([str(BIN / "nbody_mgpu"), "-t", str(threads), "-n", "4096", "-i", str(preset), "-l", str(rec["steps"]),
             "-O", str(out)], tmp_path)
        assert hashlib.sha256(out.read_bytes()).hexdigest() == rec["sha256"], (threads, preset)


def test_nbody_main_rejects_bad_options(tmp_path):
    p = subprocess.run([str(BIN / "nbody_main"), "-t", "2"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 1 and "require one thread" in p.stderr
