"""CPU tests: the oracle against the reference's golden vectors, the reference
itself (when its sources are present), and the mathematics.

No GPU needed.  The oracle is the checker for every GPU test, so each of its
functions is pinned here by something other than a retyped formula.
"""
import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
SEQ_SHA = "1484534f18e875397249e74b1a319161a7b0d70307ee4dffc1ea1683082ce1d3"


def seq_golden_run(tmp_path):
    """benchapps/nbody/seq/nbody_main.cc run as -n 1024 -l 100 (FP32 build)."""
    n = 1024
    prm = O.demo_params(0, n)
    pos, vel = O.randomize(1, n, prm["clusterScale"], prm["velocityScale"], dtype=np.float32)
    eps2 = np.float32(prm["softening"]) * np.float32(prm["softening"])
    p1, _ = O.run_f32(pos, vel, 100, eps2, prm["damping"], prm["timestep"])
    out = tmp_path / "out.txt"
    O.write_frames(out, np.stack([pos, p1]), prm["timestep"])
    return out.read_bytes(), pos


def test_oracle_reproduces_seq_golden_file(tmp_path):
    data, _ = seq_golden_run(tmp_path)
    assert hashlib.sha256(data).hexdigest() == SEQ_SHA
    assert data == (GOLDEN / "seq_nbody_outpos.txt").read_bytes()


def test_golden_frame0_is_the_shell_generator(tmp_path):
    """Frame 0 of the golden file is the SHELL initial condition at %.5f."""
    lines = (GOLDEN / "seq_nbody_outpos.txt").read_text().splitlines()
    assert lines[0] == "0.000000"
    prm = O.demo_params(0, 1024)
    pos, _ = O.randomize(1, 1024, prm["clusterScale"], prm["velocityScale"])
    for j in (0, 1, 511, 1023):
        want = " ".join(f"{v:.5f}" for v in pos[j].astype(np.float64))
        assert lines[1 + j] == want


def test_cuda_fixture_envelope_against_seq(tmp_path):
    """The reference's own GPU-vs-CPU spread on the golden run (the envelope a
    fast FP32 path is held to on this chaotic 100-step case)."""
    def frame1(p):
        txt = p.read_text().splitlines()
        return np.array([[float(x) for x in l.split()] for l in txt[1026:2050]])
    s = frame1(GOLDEN / "seq_nbody_outpos.txt")
    c = frame1(GOLDEN / "cuda_nbody_outpos.txt")
    d = np.abs(s - c)[:, :3]
    assert d.max() == pytest.approx(0.0591, abs=5e-4)
    assert np.median(d) < 1e-3


@pytest.mark.skipif(not O.ref_available(), reason="reference sources not built here")
@pytest.mark.parametrize("config", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_matches_reference_build(config, dtype):
    """Bit parity with the reference's own seq code compiled from its sources."""
    n = 512
    R = O.ref_lib()
    pos = np.zeros((n, 4), dtype); vel = np.zeros((n, 4), dtype); col = np.zeros((n, 4), np.float32)
    R.ref_srand(1)
    f = R.ref_randomize_f32 if dtype == np.float32 else R.ref_randomize_f64
    import ctypes as C
    f(config, C.c_void_p(pos.ctypes.data), C.c_void_p(vel.ctypes.data), C.c_void_p(col.ctypes.data),
      C.c_float(1.52), C.c_float(2.0), n, 1)
    op, ov = O.randomize(config, n, 1.52, 2.0, dtype=dtype)
    assert np.array_equal(pos, op) and np.array_equal(vel, ov)
    rp, rv = pos.copy(), vel.copy()
    if dtype == np.float32:
        R.ref_run_f32(n, 7, C.c_void_p(rp.ctypes.data), C.c_void_p(rv.ctypes.data), C.c_float(0.1),
                      C.c_float(0.995), C.c_float(0.016))
        eps2 = np.float32(0.1) * np.float32(0.1)
        gp, gv = O.run_f32(pos, vel, 7, eps2, np.float32(0.995), np.float32(0.016))
    else:
        R.ref_run_f64(n, 7, C.c_void_p(rp.ctypes.data), C.c_void_p(rv.ctypes.data), C.c_double(0.1),
                      C.c_double(0.995), C.c_double(0.016))
        gp, gv = O.run_f64(pos, vel, 7, 0.1 * 0.1, 0.995, 0.016)
    assert np.array_equal(rp, gp) and np.array_equal(rv, gv)


PRESETS = json.loads((Path(__file__).resolve().parent / "golden" / "preset_frames_fp64.json").read_text())


def test_preset_table_reproduces_reference_driver_frames(tmp_path):
    """Every preset (-i 0..6) at N = 1024 ... 33024 (each N branch of the
    cluster / velocity scales, and one N above them): the oracle's preset
    table, SHELL generator, FP64 steps and frame writer reproduce, byte for
    byte, the frame files of the reference's own sequential driver
    (tests/golden/preset_frames_fp64.json, made by make_preset_fixtures.py)."""
    for rec in PRESETS["frames"]:
        n, preset, steps = rec["n"], rec["preset"], rec["steps"]
        prm = O.demo_params(preset, n)
        pos, vel = O.randomize(1, n, prm["clusterScale"], prm["velocityScale"], dtype=np.float64)
        eps2 = float(prm["softening"]) * float(prm["softening"])
        p2, _ = O.run_f64(pos, vel, steps, eps2, float(prm["damping"]), float(prm["timestep"]), nthreads=8)
        path = tmp_path / f"p{preset}_{n}.txt"
        O.write_frames(path, np.stack([pos, p2]), prm["timestep"])
        data = path.read_bytes()
        assert hashlib.sha256(data).hexdigest() == rec["sha256"], (preset, n)


def test_threaded_rows_equal_serial():
    rng = np.random.default_rng(3)
    pos = np.zeros((1003, 4), np.float32)
    pos[:, :3] = rng.uniform(-1, 1, (1003, 3)); pos[:, 3] = 1 / 1003
    eps2 = np.float32(1e-4)
    full = O.gravitation_f32(pos, eps2)
    for t in (1, 3, 8):
        assert np.array_equal(O.gravitation_rows_f32(pos, eps2, 0, 1003, t), full)
    part = O.gravitation_rows_f32(pos, eps2, 100, 200, 4)
    assert np.array_equal(part[100:200], full[100:200])
    assert not part[:100].any()


# ---- FP64 physical-mode oracle ------------------------------------------------------
def test_accel_rows_f64_equal_full_field():
    """The threaded row-subset FP64 field (the yardstick of the full-size GPU
    tests) equals the serial full field bit for bit, for any row list and
    thread count; out-of-range rows are refused."""
    rng = np.random.default_rng(11)
    n = 777
    pos = np.zeros((n, 4)); pos[:, :3] = rng.normal(size=(n, 3)); pos[:, 3] = rng.uniform(0.5, 2, n) / n
    full = O.accel_f64(pos, 1e-4)
    rows = rng.choice(n, 97, replace=False)
    for t in (1, 3, 8, 200):
        assert np.array_equal(O.accel_rows_f64(pos, 1e-4, rows, t), full[rows])
    assert np.array_equal(O.accel_rows_f64(pos, 1e-4, np.arange(n), 5), full)
    with pytest.raises(ValueError):
        O.accel_rows_f64(pos, 1e-4, [0, n], 2)


def test_accel_f64_against_long_double_brute_force():
    rng = np.random.default_rng(0)
    n = 200
    pos = np.zeros((n, 4)); pos[:, :3] = rng.normal(size=(n, 3)); pos[:, 3] = rng.uniform(0.5, 2, n) / n
    a = O.accel_f64(pos, 1e-4)
    b = O.accel_ld(pos, 1e-4)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


def test_two_body_closed_form():
    pos = np.array([[0.0, 0.0, 0.0, 2.0], [0.3, -0.4, 1.2, 0.5]])
    eps2 = 0.01
    a = O.accel_f64(pos, eps2)
    r = pos[1, :3] - pos[0, :3]
    d3 = (r @ r + eps2) ** 1.5
    assert np.allclose(a[0], 0.5 * r / d3, rtol=1e-15, atol=0)
    assert np.allclose(a[1], -2.0 * r / d3, rtol=1e-15, atol=0)
    # the reference's FP32 pair on the same bodies (self pair contributes exactly 0)
    f = O.gravitation_f32(pos.astype(np.float32), np.float32(eps2))
    assert np.allclose(f[0], 0.5 * r / d3, rtol=1e-6)


def test_field_invariants():
    rng = np.random.default_rng(1)
    n = 300
    pos = np.zeros((n, 4)); pos[:, :3] = rng.uniform(-1, 1, (n, 3)); pos[:, 3] = rng.uniform(0.5, 2, n)
    eps2 = 1e-3
    a = O.accel_f64(pos, eps2)
    # momentum: sum m_i a_i = 0
    assert np.abs((pos[:, 3:4] * a).sum(0)).max() <= 1e-12 * np.abs(pos[:, 3:4] * a).sum()
    # translation invariance
    sh = pos.copy(); sh[:, :3] += np.array([0.25, -0.5, 0.125])
    assert np.abs(O.accel_f64(sh, eps2) - a).max() <= 1e-12 * np.abs(a).max()
    # scaling: a(lx, l eps) = a(x, eps) / l^2
    lam = 2.0
    sc = pos.copy(); sc[:, :3] *= lam
    a_sc = O.accel_f64(sc, eps2 * lam * lam)
    assert np.abs(a_sc - a / lam ** 2).max() <= 1e-13 * np.abs(a).max()
    # linearity in the source masses
    p2 = pos.copy(); p2[:, 3] *= 3.0
    assert np.abs(O.accel_f64(p2, eps2) - 3.0 * a).max() <= 1e-13 * np.abs(3.0 * a).max()


def test_symmetric_shell_centre_field_vanishes():
    # bodies at +-e_k pairs: the field at the centre is zero by symmetry
    pts = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [0, 0, 0]]
    pos = np.array([[*p, 1.0] for p in pts], float)
    a = O.accel_f64(pos, 1e-2)
    assert np.abs(a[-1]).max() < 1e-16


def test_energy_two_body_closed_form():
    pos = np.array([[0.0, 0.0, 0.0, 2.0], [1.0, 2.0, 2.0, 0.5]])
    vel = np.array([[0.1, 0.0, 0.0, 1.0], [0.0, -0.2, 0.3, 1.0]])
    k, w, p = O.energy_f64(pos, vel, 0.0)
    assert k == pytest.approx(0.5 * 2 * 0.01 + 0.5 * 0.5 * 0.13, rel=1e-15)
    assert w == pytest.approx(-2 * 0.5 / 3.0, rel=1e-15)
    assert np.allclose(p, [0.2, -0.1, 0.15], rtol=1e-15)


def test_step_f64_leapfrog_energy_is_bounded():
    """Kick-drift is symplectic: on a circular two-body orbit (period 4.44) the
    energy error oscillates at O(dt^2) with no secular drift over two orbits."""
    pos = np.array([[-0.5, 0, 0, 1.0], [0.5, 0, 0, 1.0]])
    v = math.sqrt(0.5)  # relative speed sqrt(G(m1+m2)/r) = sqrt(2), shared equally
    vel = np.array([[0, -v, 0, 1.0], [0, v, 0, 1.0]])
    k0, w0, _ = O.energy_f64(pos, vel, 0.0)
    e0 = k0 + w0
    assert e0 == pytest.approx(-0.5, rel=1e-15)
    p, vv = pos.copy(), vel.copy()
    errs = []
    for _ in range(40):
        p, vv = O.steps_f64(p, vv, 250, 0.0, 1e-3)
        k, w, _ = O.energy_f64(p, vv, 0.0)
        errs.append(abs(k + w - e0) / abs(e0))
    assert max(errs) < 1e-5
    assert max(errs[20:]) <= 1.05 * max(errs[:20])
    # separation stays 1 up to the O(dt) eccentricity of starting kick-drift
    # from synchronised velocities
    assert abs(np.linalg.norm(p[1, :3] - p[0, :3]) - 1.0) < 2e-3


# ---- adjoint (reverse mode) of the oracle --------------------------------------------
@pytest.mark.parametrize("kick,damping", [(1, 1.0), (0, 0.99)])
def test_oracle_adjoint_matches_central_differences(kick, damping):
    rng = np.random.default_rng(5)
    n, steps, eps2, dt = 6, 4, 0.05, 0.02
    pos = np.zeros((n, 4)); pos[:, :3] = rng.uniform(-1, 1, (n, 3)); pos[:, 3] = rng.uniform(0.5, 2, n)
    vel = np.zeros((n, 4)); vel[:, :3] = rng.normal(0, 0.1, (n, 3)); vel[:, 3] = 1
    tgt = rng.normal(0, 0.5, (n, 3))
    L, dx, dv, dm = O.adjoint_f64(pos, vel, tgt, steps, eps2, dt, damping, kick)
    assert L == pytest.approx(O.loss_f64(pos, vel, tgt, steps, eps2, dt, damping, kick), rel=1e-15)
    h = 1e-6
    for (arr, col, grad) in ((pos, range(3), dx), (vel, range(3), dv), (pos, [3], dm[:, None])):
        for i in range(n):
            for ci, c in enumerate(col):
                a = arr.copy(); a[i, c] += h
                b = arr.copy(); b[i, c] -= h
                if arr is pos:
                    fp = O.loss_f64(a, vel, tgt, steps, eps2, dt, damping, kick)
                    fm = O.loss_f64(b, vel, tgt, steps, eps2, dt, damping, kick)
                else:
                    fp = O.loss_f64(pos, a, tgt, steps, eps2, dt, damping, kick)
                    fm = O.loss_f64(pos, b, tgt, steps, eps2, dt, damping, kick)
                fd = (fp - fm) / (2 * h)
                g = grad[i, ci]
                assert abs(g - fd) <= 1e-7 * max(1.0, abs(fd)), (i, c, g, fd)


def test_oracle_adjoint_one_step_two_body_closed_form():
    """One step, two bodies, physical kick, loss on body 0 only through the
    drift: dL/dv0_x for L = (1/2)|x0' - t0|^2 + (1/2)|x1' - t1|^2."""
    pos = np.array([[0.0, 0, 0, 1.0], [1.0, 0, 0, 2.0]])
    vel = np.array([[0.0, 0, 0, 1.0], [0.0, 0, 0, 1.0]])
    tgt = np.zeros((2, 3))
    dt, eps2 = 0.1, 0.0
    L, dx, dv, dm = O.adjoint_f64(pos, vel, tgt, 1, eps2, dt, 1.0, 1)
    # x0' = x0 + dt^2 a0, a0 = m1 (x1-x0)/|x1-x0|^3 = (2, 0, 0)
    x0n = np.array([0.0 + dt * dt * 2.0, 0, 0])
    x1n = np.array([1.0 - dt * dt * 1.0, 0, 0])
    assert L == pytest.approx((x0n @ x0n + x1n @ x1n) / 2, rel=1e-14)
    # dL/dv0 = dt * dL/dx0' = dt * x0n (n = 2 -> 2/n = 1)
    assert dv[0] == pytest.approx(dt * x0n, rel=1e-14)
    # dL/dm1: a0_x = m1/r^2 (r=1) -> dx0'_x/dm1 = dt^2; a1 does not depend on m1
    assert dm[1] == pytest.approx(x0n[0] * dt * dt, rel=1e-12)
