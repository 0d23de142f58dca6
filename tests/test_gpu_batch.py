"""GPU tests of the batched systems (persistent kernels, HIP graph, per-system baseline).

* each system of a batch equals the same system run alone in a batch of one
  with the same kernel configuration, bit for bit;
* every replay of the captured graph equals the same kernels launched eagerly,
  bit for bit, with eager work and synchronisation between replays;
* the per-system baseline and the FP64 oracle agree within tolerance;
* the batched adjoint agrees with the single-system adjoint.
"""
import sys

import numpy as np
import pytest

import utc_amd as U
from nbcases import batch256_sizes, mass_pattern, relnorm
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_systems(sizes, seed0=100):
    pos, vel = [], []
    for k, n in enumerate(sizes):
        p, v = U.generate(U.GEN_PLUMMER, n, seed0 + k)
        pos.append(p)
        vel.append(v)
    return np.concatenate(pos), np.concatenate(vel)


def slices(sizes):
    o = np.concatenate([[0], np.cumsum(sizes)])
    return [slice(o[k], o[k + 1]) for k in range(len(sizes))]


SIZES = [256, 300, 1000, 64, 1, 2048, 777, 4096, 130, 512]


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5])
def test_batched_equals_single_system_batches(tile):
    """Tiles 1..3: one-pass kernels, 4 and 5: symmetric kernels (each system's
    partial sums and their reduction order depend on that system alone)."""
    eps2, dt, steps = 1e-4, 1e-3, 4
    pos, vel = make_systems(SIZES)
    say(f"case: batch vs singletons tile={tile}")
    with U.Batch(SIZES, eps2, dt, tile=tile) as b:
        b.set_state(pos, vel)
        b.run(steps, mode=2)
        bp, bv = b.get_state()
    for k, sl in enumerate(slices(SIZES)):
        with U.Batch([SIZES[k]], eps2, dt, tile=tile) as one:
            one.set_state(pos[sl], vel[sl])
            one.run(steps, mode=2)
            p1, v1 = one.get_state()
        assert np.array_equal(bp[sl], p1), k
        assert np.array_equal(bv[sl], v1), k


@pytest.mark.parametrize("tile", [2, 4])
def test_graph_replays_equal_eager(tile):
    eps2, dt, steps = 1e-4, 1e-3, 5
    pos, vel = make_systems(SIZES, 7)
    say(f"case: graph replays vs eager tile={tile}")
    with U.Batch(SIZES, eps2, dt, tile=tile) as g, U.Batch(SIZES, eps2, dt, tile=tile) as e:
        g.set_state(pos, vel)
        e.set_state(pos, vel)
        for rep in range(3):
            g.run(steps, mode=0)
            e.run(steps, mode=2)
            gp, gv = g.get_state()
            ep, ev = e.get_state()
            assert np.array_equal(gp, ep), rep
            assert np.array_equal(gv, ev), rep


@pytest.mark.parametrize("tile", [0, 1, 4, 5])
def test_batched_vs_per_system_and_oracle(tile):
    eps2, dt, steps = 1e-4, 1e-3, 4
    sizes = [256, 1024, 333, 2048]
    pos, vel = make_systems(sizes, 21)
    say(f"case: batch vs per-system vs oracle tile={tile}")
    with U.Batch(sizes, eps2, dt, tile=tile) as b:
        b.set_state(pos, vel)
        b.run(steps, mode=0)
        gp, _ = b.get_state()
        b.set_state(pos, vel)
        b.run(steps, mode=1)
        pp, _ = b.get_state()
    for sl in slices(sizes):
        p64, _ = O.steps_f64(pos[sl], vel[sl], steps, eps2, dt)
        assert relnorm(gp[sl, :3], p64[:, :3]) <= 1e-9
        assert relnorm(pp[sl, :3], p64[:, :3]) <= 1e-9


@pytest.mark.parametrize("tile", [0, 2, 4, 5])
def test_batched_adjoint_matches_single_system(tile):
    eps2, dt, steps, ck = 1e-4, 1e-3, 6, 4
    sizes = [256, 700, 1500, 64]
    pos, vel = make_systems(sizes, 33)
    tgt = U.generate(U.GEN_TARGETS, sum(sizes), 5, 0.5)
    say(f"case: batched adjoint tile={tile}")
    with U.Batch(sizes, eps2, dt, tile=tile) as b:
        b.set_state(pos, vel)
        ms, losses, dx, dv, dm = b.run(steps, mode=0, ckpt_every=ck, target3=tgt)
        b.set_state(pos, vel)
        ms1, losses1, dx1, dv1, dm1 = b.run(steps, mode=1, ckpt_every=ck, target3=tgt)
    for k, sl in enumerate(slices(sizes)):
        with U.System(sizes[k], U.PREC_FP64, eps2, dt) as s:
            s.set_state(pos[sl], vel[sl])
            L, gx, gv, gm = s.adjoint(steps, ck, tgt[sl])
        assert losses[k] == pytest.approx(L, rel=1e-9)
        assert losses1[k] == pytest.approx(L, rel=1e-9)
        for mine in ((dx, dv, dm), (dx1, dv1, dm1)):
            assert relnorm(mine[0][sl], gx) <= 1e-5
            assert relnorm(mine[1][sl], gv) <= 1e-5
            assert relnorm(mine[2][sl], gm) <= 1e-5


@pytest.mark.parametrize("tile", [4, 5])
def test_batched_symmetric_mass_patterns(tile):
    """Systems with one mass, a few heavy bodies and log-uniform masses in one
    batch: the symmetric kernels' mass-free and general pair forms side by
    side, forward against the FP64 oracle and adjoint against the FP64 system."""
    eps2, dt, steps, ck = 1e-4, 1e-3, 4, 2
    sizes = [1000, 2048, 777, 1500]
    pos, vel = make_systems(sizes, 45)
    sl = slices(sizes)
    for k, pattern in enumerate(["equal", "few_heavy", "loguniform", "one_heavy"]):
        pos[sl[k]] = mass_pattern(pos[sl[k]], pattern, seed=k)
    tgt = U.generate(U.GEN_TARGETS, sum(sizes), 9, 0.5)
    say(f"case: batched mass patterns tile={tile}")
    with U.Batch(sizes, eps2, dt, tile=tile) as b:
        b.set_state(pos, vel)
        b.run(steps, mode=0)
        gp, _ = b.get_state()
        b.set_state(pos, vel)
        _, losses, dx, dv, dm = b.run(steps, mode=0, ckpt_every=ck, target3=tgt)
    for k in range(len(sizes)):
        p64, _ = O.steps_f64(pos[sl[k]], vel[sl[k]], steps, eps2, dt)
        assert relnorm(gp[sl[k], :3], p64[:, :3]) <= 1e-9, k
        with U.System(sizes[k], U.PREC_FP64, eps2, dt) as s:
            s.set_state(pos[sl[k]], vel[sl[k]])
            L, gx, gv, gm = s.adjoint(steps, ck, tgt[sl[k]])
        assert losses[k] == pytest.approx(L, rel=1e-9), k
        assert relnorm(dx[sl[k]], gx) <= 1e-5, k
        assert relnorm(dv[sl[k]], gv) <= 1e-5, k
        assert relnorm(dm[sl[k]], gm) <= 1e-5, k


def test_batch256_full_graph_vs_per_system():
    """256 Plummer systems at full size (sizes log-uniform in
    [256, 8192]), 20 steps forward and forward + adjoint.  The tuned graph
    of batched kernels against per-system launches of the single-system
    kernels: positions within 1e-9 (both FP64 state, FP32 pairs), losses
    within 1e-9 and gradients within 1e-5 relative, per system."""
    sizes = batch256_sizes(256, 0)
    pos, vel = make_systems(sizes, 1000)
    tgt = U.generate(U.GEN_TARGETS, sum(sizes), 77, 0.5)
    say(f"case: 256-system full batch, {len(sizes)} systems, {sum(sizes)} bodies")
    with U.Batch(sizes, 1e-4, 1e-3, tile=0) as bg, U.Batch(sizes, 1e-4, 1e-3, tile=1) as bp:
        bg.set_state(pos, vel)
        bg.run(20, mode=0)
        gp, _ = bg.get_state()
        say(f"  graph kernel: {bg.kernel_name()}")
        bp.set_state(pos, vel)
        bp.run(20, mode=1)
        pp, _ = bp.get_state()
        bg.set_state(pos, vel)
        _, lg, gx, gv, gm = bg.run(20, mode=0, ckpt_every=10, target3=tgt)
        bp.set_state(pos, vel)
        _, lp, px, pv, pm = bp.run(20, mode=1, ckpt_every=10, target3=tgt)
    for k, sl in enumerate(slices(sizes)):
        assert relnorm(gp[sl, :3], pp[sl, :3]) <= 1e-9, k
        assert lg[k] == pytest.approx(lp[k], rel=1e-9), k
        assert relnorm(gx[sl], px[sl]) <= 1e-5, k
        assert relnorm(gv[sl], pv[sl]) <= 1e-5, k
        assert relnorm(gm[sl], pm[sl]) <= 1e-5, k


def test_batch_rejects_bad_arguments():
    with pytest.raises(U.UtcnbError):
        U.Batch([16, 0, 5], 1e-4, 1e-3)
    with pytest.raises(U.UtcnbError):
        U.Batch([16], 0.0, 1e-3)
    with pytest.raises(U.UtcnbError):
        U.Batch([16], 1e-4, 1e-3, tile=6)
    with U.Batch([100], 1e-4, 1e-3, kick=U.KICK_REFERENCE) as b:
        p, v = make_systems([100])
        b.set_state(p, v)
        with pytest.raises(U.UtcnbError):
            b.run(2, mode=0, ckpt_every=1, target3=np.zeros((100, 3)))


@pytest.mark.parametrize("tile", [2, 4])
def test_graph_recaptured_after_adjoint_workspace_grows(tile):
    """A mode 0 adjoint run captures a graph over the adjoint workspace; a
    later run that needs more checkpoints or a longer segment reallocates that
    workspace.  The next mode 0 run with the first run's arguments must
    capture again (not replay launches on the freed buffers) and give the
    same bits as a fresh batch."""
    eps2, dt = 1e-4, 1e-3
    sizes = [256, 700, 1500, 64]
    pos, vel = make_systems(sizes, 61)
    tgt = U.generate(U.GEN_TARGETS, sum(sizes), 13, 0.5)
    say(f"case: graph recapture after workspace growth tile={tile}")
    with U.Batch(sizes, eps2, dt, tile=tile) as fresh:
        fresh.set_state(pos, vel)
        ref = fresh.run(20, mode=0, ckpt_every=10, target3=tgt)[1:]
        ref_state = fresh.get_state()
    with U.Batch(sizes, eps2, dt, tile=tile) as b:
        b.set_state(pos, vel)
        b.run(20, mode=0, ckpt_every=10, target3=tgt)       # 2 checkpoints, segment 10
        b.set_state(pos, vel)
        b.run(30, mode=2, ckpt_every=15, target3=tgt)       # segment 15: workspace grows
        b.set_state(pos, vel)
        b.run(45, mode=1, ckpt_every=15, target3=tgt)       # 3 checkpoints: grows again
        b.set_state(pos, vel)
        got = b.run(20, mode=0, ckpt_every=10, target3=tgt)[1:]
        got_state = b.get_state()
    for r, g in zip(ref, got):
        assert np.array_equal(r, g)
    assert np.array_equal(ref_state[0], got_state[0])
    assert np.array_equal(ref_state[1], got_state[1])
