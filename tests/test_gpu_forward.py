"""GPU parity of the forward path (force + damped kick-drift) against the oracle.

Reference arithmetic (UTCNB_PREC_REFERENCE) must be bit-identical to the
reference's CPU path; the fast FP32 / mixed / FP64 paths must meet the stated
tolerances against the FP64 oracle.
"""
import hashlib
import re
import sys
from pathlib import Path

import numpy as np
import pytest

import utc_amd as U
from nbcases import cube4096, plummer65536, plummer262144, mass_pattern, plummer, relnorm
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
SEQ_SHA = "1484534f18e875397249e74b1a319161a7b0d70307ee4dffc1ea1683082ce1d3"

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fixture_run_frames(dtype):
    """benchapps/nbody/seq/nbody_main.cc as run for the golden file:
    -n 1024 -l 100, preset 0 (dt 0.016, softening 0.1, damping 1), SHELL."""
    n = 1024
    prm = O.demo_params(0, n)
    pos, vel = O.randomize(1, n, prm["clusterScale"], prm["velocityScale"], dtype=dtype)
    T = np.float32 if dtype == np.float32 else np.float64
    eps2 = T(prm["softening"]) * T(prm["softening"])
    return n, prm, pos, vel, eps2


def test_reference_f32_reproduces_golden_file(tmp_path):
    n, prm, pos, vel, eps2 = fixture_run_frames(np.float32)
    say(f"case: REFERENCE f32 golden run n={n} steps=100")
    with U.System(n, U.PREC_REFERENCE, float(eps2), float(prm["timestep"]), float(prm["damping"]),
                  kick=U.KICK_REFERENCE, state_dtype=U.F32) as s:
        s.set_state(pos, vel)
        s.step(100)
        p1, _ = s.get_state(np.float32)
    frames = np.stack([pos, p1])
    out = tmp_path / "out.txt"
    O.write_frames(out, frames, prm["timestep"])
    data = out.read_bytes()
    assert data == (GOLDEN / "seq_nbody_outpos.txt").read_bytes()
    assert hashlib.sha256(data).hexdigest() == SEQ_SHA


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reference_cube4096_bit_exact(dtype):
    c = cube4096()
    T = dtype
    pos = c["pos"].astype(T)
    vel = c["vel"].astype(T)
    eps2 = T(c["eps"]) * T(c["eps"])
    say(f"case: REFERENCE {np.dtype(T).name} cube4096 n={len(pos)}")
    with U.System(len(pos), U.PREC_REFERENCE, float(eps2), float(T(c["dt"])), 1.0,
                  kick=U.KICK_REFERENCE, state_dtype=U.F32 if T == np.float32 else U.F64) as s:
        s.set_state(pos, vel)
        s.step(c["steps"])
        gp, gv = s.get_state(T)
    if T == np.float32:
        op, ov = O.run_f32(pos, vel, c["steps"], eps2, 1.0, T(c["dt"]), nthreads=8)
    else:
        op, ov = O.run_f64(pos, vel, c["steps"], eps2, 1.0, c["dt"])
    assert np.array_equal(gp, op), f"max diff {np.abs(gp - op).max()}"
    assert np.array_equal(gv, ov)


def test_reference_inverse_distance_exhaustive():
    """The reference kernels' inverse distance for eps2 >= 2^-96 (ref_inv.h):
    every float d2 in [2^-96, inf] and NaN against the compiler's correctly
    rounded (float)(1.0 / (double)sqrtf(d2)) on the device, bit for bit; the
    FP64 one against 1.0 / sqrt(d2) on 2^32 random d2 in [2^-600, 2^600]."""
    import json
    import subprocess
    exe = Path(__file__).resolve().parent / "device" / "_bin" / "probe_invdist"
    assert exe.exists(), "build first: python -m utc_amd.build"
    say("case: probe_invdist (all floats 2^-96..inf)")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    say(f"  {out}")
    assert r.returncode == 0
    assert out["mismatches"] == 0
    assert out["checked"] == 0x7F800001 - 0x0F800000 + 1
    # the margins the correction steps start from: v_sqrt_f32 and v_rcp_f32
    # within 1 ulp of the correctly rounded root and reciprocal
    assert out["sqrt_f32_max_ulps"] <= 1
    assert out["rcp_f32_max_ulps"] <= 1
    # FP64 (ref_inv_dist64): 2^32 random d2 over [2^-600, 2^600] and both ends
    assert out["f64_checked"] == 2 ** 32 + 2
    assert out["f64_mismatches"] == 0
    # and m / sqrt(d2) (the energy kernel) on 2^31 random (m, d2)
    assert out["f64div_checked"] == 2 ** 31
    assert out["f64div_mismatches"] == 0


@pytest.mark.parametrize("eps2", [1e-30, 1e-4])
def test_reference_f32_far_and_tiny_softening_bit_exact(eps2):
    """eps2 below 2^-96 takes the generic sqrtf and division; pairs 1e20 apart
    overflow d2 to inf, where the reference's inverse distance is 0."""
    rng = np.random.default_rng(3)
    n = 1000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    pos[:, 3] = rng.uniform(0.5, 2, n).astype(np.float32)
    pos[:10, 0] += np.float32(1e20)   # far cluster: d2 = inf against the rest
    pos[10:20, 1] -= np.float32(3e19)
    say(f"case: REFERENCE f32 field eps2={eps2} with overflowing pairs")
    e2 = np.float32(eps2)
    with U.System(n, U.PREC_REFERENCE, float(e2), 1e-3, 1.0, kick=U.KICK_REFERENCE,
                  state_dtype=U.F32) as s:
        s.set_state(pos, np.zeros_like(pos))
        a = s.accel()
    f = O.gravitation_f32(pos, e2)
    assert np.array_equal(a.astype(np.float32), f, equal_nan=True)


@pytest.mark.parametrize("n", [1, 3, 17, 1003])
def test_reference_ragged_sizes_bit_exact(n):
    """Sizes that leave remainders after the kernels' 16- and 4-source trips
    (FP32: 1003 = 62 x 16 + 2 x 4 + 3), both state types, 3 steps, against the
    oracle's sequential loop over every source."""
    c = cube4096(n=n)
    say(f"case: REFERENCE ragged n={n}")
    for T, sd in ((np.float32, U.F32), (np.float64, U.F64)):
        pos = c["pos"].astype(T)
        vel = c["vel"].astype(T)
        eps2 = T(c["eps"]) * T(c["eps"])
        with U.System(n, U.PREC_REFERENCE, float(eps2), float(T(c["dt"])), 1.0,
                      kick=U.KICK_REFERENCE, state_dtype=sd) as s:
            s.set_state(pos, vel)
            s.step(3)
            gp, gv = s.get_state(T)
        if T == np.float32:
            op, ov = O.run_f32(pos, vel, 3, eps2, 1.0, T(c["dt"]))
        else:
            op, ov = O.run_f64(pos, vel, 3, eps2, 1.0, c["dt"])
        assert np.array_equal(gp, op), (n, T)
        assert np.array_equal(gv, ov), (n, T)


@pytest.mark.parametrize("far", [1e100, 1e160])
def test_reference_f64_far_pairs_bit_exact(far):
    """FP64 reference arithmetic with pairs whose d2 leaves [2^-600, 2^600]
    (1e100 apart: d2 = 1e200; 1e160 apart: d2 overflows to inf, inverse
    distance 0): those groups of four sources take the generic sqrt and
    division, the rest the short sequence; the field equals the oracle's."""
    rng = np.random.default_rng(4)
    n = 1000
    pos = np.zeros((n, 4))
    pos[:, :3] = rng.uniform(-1, 1, (n, 3))
    pos[:, 3] = rng.uniform(0.5, 2, n)
    pos[:7, 2] += far
    say(f"case: REFERENCE f64 field with pairs {far:g} apart")
    with U.System(n, U.PREC_REFERENCE, 1e-4, 1e-3, 1.0, kick=U.KICK_REFERENCE, state_dtype=U.F64) as s:
        s.set_state(pos, np.zeros_like(pos))
        a = s.accel()
    f = O.gravitation_f64(pos, 1e-4)
    assert np.array_equal(a, f, equal_nan=True)


def test_reference_f32_field_matches_oracle_gravitation():
    n, prm, pos, vel, eps2 = fixture_run_frames(np.float32)
    say("case: REFERENCE f32 field n=1024")
    with U.System(n, U.PREC_REFERENCE, float(eps2), 0.016, 1.0, kick=U.KICK_REFERENCE,
                  state_dtype=U.F32) as s:
        s.set_state(pos, vel)
        a = s.accel()
    f = O.gravitation_f32(pos, eps2)
    assert np.array_equal(a.astype(np.float32), f)


def test_kernel_level_entry_point_matches_oracle():
    """utcnb_integrate_f32 on device pointers, like the reference's kernel launch."""
    import torch
    c = cube4096(n=2048)
    pos = c["pos"].astype(np.float32)
    vel = c["vel"].astype(np.float32)
    eps2 = np.float32(c["eps"]) * np.float32(c["eps"])
    dt = np.float32(c["dt"])
    say("case: utcnb_integrate_f32 n=2048")
    dev = torch.device("cuda:0")
    p_old = torch.from_numpy(pos).to(dev)
    p_new = torch.empty_like(p_old)
    v = torch.from_numpy(vel).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    for blocksize in (0, 64, 256, 1024):
        p_new.zero_()
        v.copy_(torch.from_numpy(vel))
        rc = U.lib().utcnb_integrate_f32(p_old.data_ptr(), p_new.data_ptr(), v.data_ptr(), len(pos),
                                         float(dt), float(eps2), 1.0, blocksize, stream)
        U.check(rc, "utcnb_integrate_f32")
        torch.cuda.synchronize()
        op, ov = O.run_f32(pos, vel, 1, eps2, 1.0, dt, nthreads=8)
        assert np.array_equal(p_new.cpu().numpy(), op)
        assert np.array_equal(v.cpu().numpy(), ov)


def test_kernel_level_entry_point_f64_matches_oracle():
    """utcnb_integrate_f64 (the reference's FTYPE = double build) on device
    pointers, a size that is not a multiple of the block size."""
    import torch
    c = cube4096(n=2050)
    pos = c["pos"].astype(np.float64)
    vel = c["vel"].astype(np.float64)
    eps2 = c["eps"] * c["eps"]
    say("case: utcnb_integrate_f64 n=2050")
    dev = torch.device("cuda:0")
    p_old = torch.from_numpy(pos).to(dev)
    p_new = torch.empty_like(p_old)
    v = torch.from_numpy(vel).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    op, ov = O.run_f64(pos, vel, 1, eps2, 1.0, c["dt"])
    for blocksize in (0, 64, 256):
        p_new.zero_()
        v.copy_(torch.from_numpy(vel))
        rc = U.lib().utcnb_integrate_f64(p_old.data_ptr(), p_new.data_ptr(), v.data_ptr(), len(pos),
                                         c["dt"], eps2, 1.0, blocksize, stream)
        U.check(rc, "utcnb_integrate_f64")
        torch.cuda.synchronize()
        assert np.array_equal(p_new.cpu().numpy(), op), blocksize
        assert np.array_equal(v.cpu().numpy(), ov), blocksize


@pytest.mark.parametrize("tile", ["default", "k2s8", "sym2r32", "sym4r32"])
def test_fp32_cube4096_within_tolerance_of_reference(tile):
    """Plain fast FP32 (rsq, FMA, uncompensated FP32 sums, FP32 state) vs the
    reference CPU path on the N=4096 cube: max |dx| <= 1e-6 and relative norm
    <= 1e-6.  The default tile at this size (one kernel per step: plain FP32
    lane sums combined in FP32), a one-pass tile and two symmetric ones
    (plain FP32 block sums and an FP32 reduction of the partials)."""
    c = cube4096()
    pos = c["pos"].astype(np.float32)
    vel = c["vel"].astype(np.float32)
    eps2 = np.float32(c["eps"]) * np.float32(c["eps"])
    dt = np.float32(c["dt"])
    op, _ = O.run_f32(pos, vel, c["steps"], eps2, 1.0, dt, nthreads=8)
    tid = 0 if tile == "default" else U.tiles().index(tile) + 1
    for flags in (0, U.FLAG_NO_NEWTON):
        say(f"case: FP32 cube4096 tile={tile} flags={flags}")
        with U.System(len(pos), U.PREC_FP32, float(eps2), float(dt), 1.0,
                      kick=U.KICK_REFERENCE, flags=flags, tile=tid) as s:
            if tile == "default":
                assert s.tile()[1] == "sm1t16w16n4k"
            s.set_state(pos, vel)
            s.step(c["steps"])
            gp, _ = s.get_state(np.float32)
        d = np.abs(gp[:, :3].astype(np.float64) - op[:, :3])
        assert d.max() <= 1e-6, d.max()
        assert relnorm(gp[:, :3], op[:, :3]) <= 1e-6
        assert np.array_equal(gp[:, 3], op[:, 3])


def test_fp64_field_matches_cpu_fp64():
    c = plummer(2048, eps=0.005)
    say("case: FP64 field n=2048")
    a64 = O.accel_f64(c["pos"], c["eps"] ** 2)
    with U.System(2048, U.PREC_FP64, c["eps"] ** 2, c["dt"]) as s:
        assert s.tile() == (0, "")  # the FP64 kernel has no tiles
        s.set_state(c["pos"], c["vel"])
        a = s.accel()
    assert relnorm(a, a64) <= 1e-14
    assert np.abs(a - a64).max() <= 1e-12 * np.abs(a64).max()


def small_cap(name):
    """Bodies a small-system tile holds: its name ends in nCAPk (CAP/1024)."""
    return int(re.search(r"n(\d+)k", name).group(1)) * 1024


@pytest.mark.parametrize("prec", [U.PREC_FP32, U.PREC_MIXED])
@pytest.mark.parametrize("tile", list(range(1, U.lib().utcnb_tile_count() + 1)))
def test_every_tile_field_accuracy(prec, tile):
    n = 3000  # not a multiple of any tile: exercises clamped targets and padding
    name = U.tiles()[tile - 1]
    if name.startswith("sm"):  # small-system tiles hold at most CAP bodies (name ends nCAPk)
        n = min(n, small_cap(name) - 47)
    c = plummer(n, eps=0.005)
    a64 = O.accel_f64(c["pos"], c["eps"] ** 2)
    say(f"case: field prec={prec} tile={tile} n={n}")
    with U.System(n, prec, c["eps"] ** 2, c["dt"], tile=tile) as s:
        s.set_state(c["pos"], c["vel"])
        a = s.accel()
    # the FP32 positions alone perturb close pairs by ~u|x|/r
    assert relnorm(a, a64) <= 2e-6


@pytest.mark.parametrize("name", [t for t in U.tiles() if t.startswith("sm")])
def test_small_tiles_reject_systems_above_their_capacity(name):
    cap = small_cap(name)
    with U.System(cap, U.PREC_MIXED, 1e-4, 1e-3, tile=U.tiles().index(name) + 1):
        pass
    with pytest.raises(U.UtcnbError, match="holds at most"):
        U.System(cap + 1, U.PREC_MIXED, 1e-4, 1e-3, tile=U.tiles().index(name) + 1)


@pytest.mark.parametrize("name", ["sm1t8w16n8k", "sm2t16w16n8k", "sym4r64x2", "sym2r128x4"])
def test_small_and_chunked_rotation_tiles_are_repeatable(name):
    """Fixed summation orders: two handles stepping the same state give
    bitwise identical trajectories (mixed and plain FP32), and the small-system
    tiles' field matches the CPU FP64 oracle at a ragged size."""
    n = 5003
    c = plummer(n, eps=0.005)
    tile = U.tiles().index(name) + 1
    say(f"case: repeatability tile={name} n={n}")
    for prec in (U.PREC_MIXED, U.PREC_FP32):
        out = []
        for _ in range(2):
            with U.System(n, prec, c["eps"] ** 2, c["dt"], tile=tile) as s:
                s.set_state(c["pos"], c["vel"])
                s.step(5)
                out.append(s.get_state())
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1]), prec
    a64 = O.accel_f64(c["pos"], c["eps"] ** 2)
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], tile=tile) as s:
        s.set_state(c["pos"], c["vel"])
        assert relnorm(s.accel(), a64) <= 2e-6


@pytest.mark.parametrize("pattern", ["equal", "one_heavy", "few_heavy", "loguniform"])
@pytest.mark.parametrize("name", ["sym2r32", "sym4r32", "sym8r32", "sym8r64", "sym8r4", "sym2r128x4", "sym4r64x2",
                                  "sym4r128x4"])
def test_symmetric_tiles_mass_patterns(name, pattern):
    n = 6000  # not a multiple of 512: the last wave block holds padding
    c = plummer(n, eps=0.005)
    pos = mass_pattern(c["pos"], pattern)
    tile = U.tiles().index(name) + 1
    a64 = O.accel_f64(pos, c["eps"] ** 2)
    say(f"case: symmetric field tile={name} masses={pattern} n={n}")
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], tile=tile) as s:
        s.set_state(pos, c["vel"])
        a = s.accel()
    assert relnorm(a, a64) <= 2e-6
    # every body, not only the norm over all bodies
    assert np.abs(a - a64).max() <= 1e-5 * np.abs(a64).max()


@pytest.mark.parametrize("n", [1, 7, 64, 1000, 4097])
def test_odd_sizes(n):
    c = plummer(n, eps=0.01, dt=1e-3, steps=3)
    say(f"case: odd size n={n}")
    p64, v64 = O.steps_f64(c["pos"], c["vel"], 3, c["eps"] ** 2, c["dt"])
    for prec in (U.PREC_FP32, U.PREC_MIXED, U.PREC_FP64):
        with U.System(n, prec, c["eps"] ** 2, c["dt"]) as s:
            s.set_state(c["pos"], c["vel"])
            s.step(3)
            gp, gv = s.get_state()
        tol = 1e-12 if prec == U.PREC_FP64 else 1e-6
        assert np.abs(gp - p64).max() <= tol * max(1.0, np.abs(p64).max()), (prec, n)


def test_mixed_and_fp64_trajectory_plummer_small():
    """Config-2 physics at N=4096: mixed path final positions vs the FP64 oracle."""
    c = plummer(4096)
    say("case: MIXED/FP64 trajectory n=4096 steps=20")
    p64, v64 = O.steps_f64(c["pos"], c["vel"], c["steps"], c["eps"] ** 2, c["dt"])
    res = {}
    for prec in (U.PREC_FP32, U.PREC_MIXED, U.PREC_FP64):
        with U.System(4096, prec, c["eps"] ** 2, c["dt"]) as s:
            s.set_state(c["pos"], c["vel"])
            s.step(c["steps"])
            gp, _ = s.get_state()
        res[prec] = relnorm(gp[:, :3], p64[:, :3])
    say(f"final-position rel err: {res}")
    assert res[U.PREC_FP64] <= 1e-13
    assert res[U.PREC_MIXED] <= 1e-8
    assert res[U.PREC_FP32] <= 1e-6


def test_diagnostics_match_oracle():
    c = plummer(2048, eps=0.01)
    say("case: diagnostics n=2048")
    K, W, P = O.energy_f64(c["pos"], c["vel"], c["eps"] ** 2)
    for prec in (U.PREC_MIXED, U.PREC_FP64):
        with U.System(2048, prec, c["eps"] ** 2, c["dt"]) as s:
            s.set_state(c["pos"], c["vel"])
            k, w, p = s.diagnostics()
        assert abs(k - K) <= 1e-13 * abs(K)
        assert abs(w - W) <= 1e-12 * abs(W)
        assert np.abs(p - P).max() <= 1e-15


@pytest.mark.parametrize("which", ["plummer65536", "plummer262144"])
def test_full_size_accuracy_against_fp64(which):
    """Configs 2 and 3 at full size (N = 65536 and 262144, Plummer): final
    positions of the mixed path (default tile, symmetric) and of plain FP32
    against the on-device FP64 path (itself equal to the CPU FP64 oracle to
    1e-14 at N = 2048, test_fp64_field_matches_cpu_fp64); target <= 1e-6
    relative.  The mixed run's total energy and momentum stay put."""
    c = plummer65536() if which == "plummer65536" else plummer262144()
    n, eps2, steps = len(c["pos"]), c["eps"] ** 2, c["steps"]
    say(f"case: {which} full size n={n} steps={steps}")
    with U.System(n, U.PREC_FP64, eps2, c["dt"]) as s64:
        s64.set_state(c["pos"], c["vel"])
        s64.step(steps)
        p64, _ = s64.get_state()
    errs = {}
    for prec in (U.PREC_MIXED, U.PREC_FP32):
        with U.System(n, prec, eps2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
            assert s.tile()[1].startswith("sym")
            s.set_state(c["pos"], c["vel"])
            k0, w0, _ = s.diagnostics()
            s.step(steps)
            gp, _ = s.get_state()
            k1, w1, mom = s.diagnostics()
        errs[prec] = relnorm(gp[:, :3], p64[:, :3])
        if prec == U.PREC_MIXED:
            drift = abs((k1 + w1) - (k0 + w0)) / abs(k0 + w0)
            say(f"  mixed: energy drift {drift:.3e}, |momentum| {np.abs(mom).max():.3e}")
            assert drift <= 1e-6
            assert np.abs(mom).max() <= 1e-8
    say(f"  final-position rel err vs FP64: mixed {errs[U.PREC_MIXED]:.3e}, fp32 {errs[U.PREC_FP32]:.3e}")
    assert errs[U.PREC_MIXED] <= 1e-6
    assert errs[U.PREC_FP32] <= 1e-6


@pytest.mark.parametrize("which", ["plummer65536", "plummer131072", "plummer262144"])
def test_full_size_fields_against_cpu_fp64_oracle(which):
    """The yardstick of the full-size tests, pinned to the CPU: at N = 65536
    (configuration 2), 131072 and 262144 (configuration 3) -- the three launch
    shapes of the FP64 kernel (4, 8 and 16 source splits) -- the FP64 field of
    512 sampled bodies against the CPU FP64 oracle (compensated sums over all N
    sources), at the initial state and after the configuration's steps; and
    the mixed (compensated) and plain FP32 fields of the same bodies against
    it.  Tolerances: FP64 1e-14 relative, mixed 2e-6, plain FP32 4e-6."""
    if which == "plummer131072":
        c = plummer(131072, eps=0.002, dt=2.5e-4, steps=10)
    else:
        c = plummer65536() if which == "plummer65536" else plummer262144()
    n, eps2 = len(c["pos"]), c["eps"] ** 2
    rows = np.sort(np.random.default_rng(n).choice(n, 512, replace=False))
    say(f"case: {which} sampled rows vs CPU FP64 oracle n={n}")
    with U.System(n, U.PREC_FP64, eps2, c["dt"]) as s64:
        s64.set_state(c["pos"], c["vel"])
        a64 = s64.accel()
        s64.step(c["steps"])
        p_end, _ = s64.get_state()
        a64_end = s64.accel()
    for tag, pos, field in (("initial", c["pos"], a64), ("final", p_end, a64_end)):
        ref = O.accel_rows_f64(pos, eps2, rows, nthreads=16)
        err = relnorm(field[rows], ref)
        say(f"  FP64 field ({tag} state) vs CPU oracle: rel {err:.3e}")
        assert err <= 1e-14, tag
        assert np.abs(field[rows] - ref).max() <= 1e-13 * np.abs(ref).max(), tag
    ref = O.accel_rows_f64(c["pos"], eps2, rows, nthreads=16)
    errs = {}
    for prec in (U.PREC_MIXED, U.PREC_FP32):
        with U.System(n, prec, eps2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
            s.set_state(c["pos"], c["vel"])
            errs[prec] = relnorm(s.accel()[rows], ref)
    say(f"  mixed field rel {errs[U.PREC_MIXED]:.3e}, plain FP32 field rel {errs[U.PREC_FP32]:.3e}")
    assert errs[U.PREC_MIXED] <= 2e-6
    assert errs[U.PREC_FP32] <= 4e-6


def test_large_n_symmetric_default_matches_fp64():
    """N = 1.5M: the symmetric kernel's partial sums (26 GB) fit the device's
    budget (a quarter of its memory, at most 64 GiB), so the default tile stays
    symmetric; its field against the on-device FP64 path."""
    n = 1_500_000
    c = plummer(n, eps=0.002, dt=2.5e-4)
    say(f"case: large N={n} default tile vs FP64 field")
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"]) as s:
        name = s.tile()[1]
        assert name.startswith("sym"), name
        s.set_state(c["pos"], c["vel"])
        a = s.accel()
    with U.System(n, U.PREC_FP64, c["eps"] ** 2, c["dt"]) as s64:
        s64.set_state(c["pos"], c["vel"])
        a64 = s64.accel()
    err = relnorm(a, a64)
    say(f"  tile {name}: accel rel err {err:.3e}")
    assert err <= 2e-6


def test_autotune_picks_a_valid_tile_and_caches_it_by_name(tmp_path):
    """utcnb_autotune times every tile for the size class (N rounded up to 1024),
    returns a tile usable in utcnb_config.tile, and keeps the result in memory
    and in the cache file (one line: arch, precision, class, tile name, ms)."""
    cache = tmp_path / "tiles.txt"
    n = 5000
    say(f"case: autotune n={n}")
    tile, ms = U.autotune(n, U.PREC_MIXED, cache)
    assert 1 <= tile <= U.lib().utcnb_tile_count() and ms > 0
    fields = cache.read_text().split()
    assert fields[0].startswith("gfx950") and fields[1] == str(U.PREC_MIXED)
    assert fields[2] == "5120" and fields[3] == U.tiles()[tile - 1]
    assert U.autotune(n, U.PREC_MIXED, cache) == (tile, ms)  # from the cache
    # a file from a build whose tile table differs: names, not ids, are kept
    cache.write_text(f"{fields[0]} {U.PREC_FP32} 5120 no-such-tile 1.0\n"
                     f"{fields[0]} {U.PREC_FP32} 6144 k1s8 0.5\n")
    t2, ms2 = U.autotune(6000, U.PREC_FP32, cache)
    assert (t2, ms2) == (U.tiles().index("k1s8") + 1, 0.5)
    c = plummer(n, eps=0.005)
    a64 = O.accel_f64(c["pos"], c["eps"] ** 2)
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], tile=tile) as s:
        s.set_state(c["pos"], c["vel"])
        a = s.accel()
    assert relnorm(a, a64) <= 2e-6


def test_long_run_mixed_tracks_fp64():
    """500 steps (half a crossing time) of a 16384-body Plummer sphere, eps = 0.01,
    dt = 1e-3: the mixed path (default symmetric tile) stays within 1e-6 of the
    FP64 path in final positions, and both conserve energy to 1e-5."""
    n, steps = 16384, 500
    c = plummer(n, eps=0.01, dt=1e-3)
    say(f"case: long run n={n} steps={steps}")
    out = {}
    for prec in (U.PREC_FP64, U.PREC_MIXED):
        with U.System(n, prec, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
            s.set_state(c["pos"], c["vel"])
            k0, w0, _ = s.diagnostics()
            s.step(steps)
            p, _ = s.get_state()
            k1, w1, _ = s.diagnostics()
        out[prec] = (p, abs((k1 + w1) - (k0 + w0)) / abs(k0 + w0))
    err = relnorm(out[U.PREC_MIXED][0][:, :3], out[U.PREC_FP64][0][:, :3])
    say(f"  final-position rel err {err:.3e}, energy drift fp64 {out[U.PREC_FP64][1]:.3e} "
        f"mixed {out[U.PREC_MIXED][1]:.3e}")
    assert err <= 1e-6
    assert out[U.PREC_FP64][1] <= 1e-5 and out[U.PREC_MIXED][1] <= 1e-5


def test_set_state_host_buffers_reusable_on_return():
    """utcnb_set_state from host memory returns after its copies: overwriting
    the caller's arrays right after the call does not change the state."""
    n = 4096
    c = plummer(n)
    pos, vel = c["pos"].copy(), c["vel"].copy()
    say("case: set_state host buffer reuse")
    for prec in (U.PREC_MIXED, U.PREC_FP32, U.PREC_FP64):
        with U.System(n, prec, c["eps"] ** 2, c["dt"]) as s:
            s.set_state(pos, vel)
            pos[:] = 7.0
            vel[:] = -7.0
            gp, gv = s.get_state()
            pos[:], vel[:] = c["pos"], c["vel"]
        tol = 0.0 if prec == U.PREC_FP64 else 1e-6
        assert np.abs(gp - c["pos"]).max() <= tol * np.abs(c["pos"]).max()
        assert np.abs(gv - c["vel"]).max() <= tol * np.abs(c["vel"]).max() + tol
