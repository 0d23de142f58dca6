"""GPU parity of the adjoint (reverse mode through force + kick-drift).

Checks: the FP64 GPU adjoint against the CPU FP64 reverse mode of the oracle
(itself pinned by finite differences in test_oracle.py); GPU gradients against
FP64 central finite differences at the config-4 size (N=16384, 50 steps,
checkpoint every 10, 64 random coordinates); the mixed (FP32-pair) adjoint
against the FP64 one; checkpoint-interval invariance; and a torch.autograd
wrapper, which exists only here in the test harness.
"""
import sys

import numpy as np
import pytest

import utc_amd as U
from nbcases import cube_logm16384, mass_pattern, relnorm
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_case(n=256, steps=12, seed=3, kick=U.KICK_PHYSICAL):
    pos, vel = U.generate(U.GEN_CUBE_LOGM, n, seed, 0.05, 0.5, 2.0)
    tgt = U.generate(U.GEN_TARGETS, n, seed + 1, 0.5)
    return pos, vel, tgt


@pytest.mark.parametrize("kick,damping", [(U.KICK_PHYSICAL, 1.0), (U.KICK_REFERENCE, 0.995)])
def test_fp64_adjoint_matches_cpu_reverse_mode(kick, damping):
    n, steps, eps2, dt = 256, 12, 1e-3, 2e-3
    pos, vel, tgt = small_case(n, steps)
    say(f"case: FP64 adjoint vs CPU oracle kick={kick} n={n}")
    L0, dx0, dv0, dm0 = O.adjoint_f64(pos, vel, tgt, steps, eps2, dt, damping, kick)
    with U.System(n, U.PREC_FP64, eps2, dt, damping=damping, kick=kick) as s:
        s.set_state(pos, vel)
        L, dx, dv, dm = s.adjoint(steps, 5, tgt)
    assert L == pytest.approx(L0, rel=1e-12)
    assert relnorm(dx, dx0) <= 1e-11
    assert relnorm(dv, dv0) <= 1e-11
    assert relnorm(dm, dm0) <= 1e-10


@pytest.mark.parametrize("prec", [U.PREC_MIXED, U.PREC_FP64])
def test_checkpoint_interval_does_not_change_gradients(prec):
    n, steps, eps2, dt = 2048, 20, 1e-4, 1e-3
    pos, vel, tgt = small_case(n, steps)
    say(f"case: checkpoint invariance prec={prec}")
    outs = []
    with U.System(n, prec, eps2, dt) as s:
        for k in (1, 7, 10, 20):  # 20, 3, 2 and 1 segments
            s.set_state(pos, vel)
            outs.append(s.adjoint(steps, k, tgt))
    for L, dx, dv, dm in outs[1:]:
        assert L == outs[0][0]
        assert np.array_equal(dx, outs[0][1])
        assert np.array_equal(dv, outs[0][2])
        assert np.array_equal(dm, outs[0][3])


def fp64_final(s, pos, vel, steps):
    s.set_state(pos, vel)
    s.step(steps)
    p, _ = s.get_state()
    return p[:, :3]


def central_difference(p_plus, p_minus, tgt, h):
    """(L(+h) - L(-h)) / 2h for L = (1/n) sum |x - t|^2, evaluated per body as
    (x+ - x-) . (x+ + x- - 2t): the same number without subtracting two
    nearly equal totals over all n bodies."""
    d = (p_plus - p_minus) * (p_plus + p_minus - 2.0 * tgt)
    return float(d.sum() / (len(tgt) * 2.0 * h))


@pytest.fixture(scope="module")
def cube_logm_grads():
    c = cube_logm16384()
    n, steps = len(c["pos"]), c["steps"]
    eps2 = c["eps"] ** 2
    say(f"case: cube_logm16384 adjoint n={n} steps={steps}")
    res = {}
    for prec in (U.PREC_FP64, U.PREC_MIXED):
        with U.System(n, prec, eps2, c["dt"], kick=c["kick"]) as s:
            s.set_state(c["pos"], c["vel"])
            res[prec] = s.adjoint(steps, c["ckpt"], c["target"])
    return c, res


def test_cube_logm16384_gradients_match_fp64_central_differences(cube_logm_grads):
    """64 random coordinates (positions, velocities, masses) of the N=16384 cube; the
    loss of each perturbed input is an FP64 forward on the GPU (itself checked
    against the CPU FP64 oracle in test_gpu_forward.py).  FP64 rounding leaves
    ~1e-13 of noise in the per-body sum of position differences whatever the
    step, so the steps are as large as the loss's curvature allows: 1e-6 for a
    position (close pairs sit at eps = 0.01), 1e-4 for a velocity and 5% of the
    body's mass for a mass; the loss is close to quadratic in v and m over 50
    steps, where central differences are exact.  Tolerance: 1e-5 relative per
    coordinate."""
    c, res = cube_logm_grads
    n, steps = len(c["pos"]), c["steps"]
    eps2 = c["eps"] ** 2
    rng = np.random.default_rng(11)
    kinds = rng.integers(0, 3, 64)       # 0: position, 1: velocity, 2: mass
    bodies = rng.integers(0, n, 64)
    comps = rng.integers(0, 3, 64)
    fds, g64, gmx = [], [], []
    with U.System(n, U.PREC_FP64, eps2, c["dt"], kick=c["kick"]) as s:
        for kind, i, comp in zip(kinds, bodies, comps):
            if kind == 2:
                comp = 3
            arr = c["vel"] if kind == 1 else c["pos"]
            h = (1e-6, 1e-4, 0.05 * c["pos"][i, 3])[kind]
            a = arr.copy(); a[i, comp] += h
            b = arr.copy(); b[i, comp] -= h
            if kind == 1:
                pp = fp64_final(s, c["pos"], a, steps)
                pm = fp64_final(s, c["pos"], b, steps)
            else:
                pp = fp64_final(s, a, c["vel"], steps)
                pm = fp64_final(s, b, c["vel"], steps)
            fds.append(central_difference(pp, pm, c["target"], h))
            for store, prec in ((g64, U.PREC_FP64), (gmx, U.PREC_MIXED)):
                L, dx, dv, dm = res[prec]
                store.append(dm[i] if kind == 2 else (dv if kind == 1 else dx)[i, comp])
    fds, g64, gmx = map(np.array, (fds, g64, gmx))
    e64 = np.abs(g64 - fds) / np.abs(fds)
    emx = np.abs(gmx - fds) / np.abs(fds)
    say(f"FD check: FP64 max rel {e64.max():.3e}, mixed max rel {emx.max():.3e}, "
        f"mixed normwise {relnorm(gmx, fds):.3e}")
    assert e64.max() <= 1e-5
    assert emx.max() <= 1e-5


def test_cube_logm16384_mixed_adjoint_close_to_fp64(cube_logm_grads):
    c, res = cube_logm_grads
    L64, dx64, dv64, dm64 = res[U.PREC_FP64]
    L, dx, dv, dm = res[U.PREC_MIXED]
    assert L == pytest.approx(L64, rel=1e-9)
    assert relnorm(dx, dx64) <= 1e-5
    assert relnorm(dv, dv64) <= 1e-5
    assert relnorm(dm, dm64) <= 1e-5


def tile_id(name):
    return U.tiles().index(name) + 1


@pytest.mark.parametrize("kick", [U.KICK_PHYSICAL, U.KICK_REFERENCE])
def test_self_pair_exact_at_tiny_softening(kick):
    """eps = 1e-6: the pair sums include j = i, where J(0) = eps^-3 I = 1e18 I
    multiplies w = m_i a_i - m_i a_i.  The kernels must make that w exactly 0
    (the oracle skips j = i); a rounding residual of m_i a_i there would swamp
    the gradient.  FP64, and mixed with the one-pass and the symmetric VJP."""
    n, steps, eps2, dt = 200, 4, 1e-12, 2e-3
    pos, vel, tgt = small_case(n, steps, seed=5)
    L0, dx0, dv0, dm0 = O.adjoint_f64(pos, vel, tgt, steps, eps2, dt, 1.0, kick)
    for prec, tile, tol in ((U.PREC_FP64, 0, 1e-11), (U.PREC_MIXED, tile_id("k1s16"), 1e-5),
                            (U.PREC_MIXED, tile_id("sym4r32"), 1e-5)):
        say(f"case: self pair eps=1e-6 kick={kick} prec={prec} tile={tile}")
        with U.System(n, prec, eps2, dt, kick=kick, tile=tile) as s:
            s.set_state(pos, vel)
            L, dx, dv, dm = s.adjoint(steps, 2, tgt)
        errs = (relnorm(dx, dx0), relnorm(dv, dv0), relnorm(dm, dm0))
        say(f"  rel err dx {errs[0]:.2e} dv {errs[1]:.2e} dm {errs[2]:.2e}")
        assert max(errs) <= tol


def test_symmetric_vjp_matches_one_pass_and_fp64():
    """A symmetric force tile also evaluates the adjoint pair sum once per
    unordered pair (sym_vjp.hip).  N = 5000 (padded to 5120, ragged last wave
    block), physical kick: gradients of the symmetric VJP against the
    one-pass VJP and the FP64 adjoint."""
    n, steps, eps2, dt = 5000, 6, 1e-4, 1e-3
    pos, vel, tgt = small_case(n, steps, seed=7)
    res = {}
    for label, prec, tile in (("fp64", U.PREC_FP64, 0), ("onepass", U.PREC_MIXED, tile_id("k1s16")),
                              ("sym", U.PREC_MIXED, tile_id("sym4r32")),
                              ("sym8", U.PREC_MIXED, tile_id("sym8r32"))):
        say(f"case: VJP {label} n={n}")
        with U.System(n, prec, eps2, dt, tile=tile) as s:
            s.set_state(pos, vel)
            res[label] = s.adjoint(steps, 3, tgt)
    for label in ("onepass", "sym", "sym8"):
        e = [relnorm(res[label][q], res["fp64"][q]) for q in (1, 2, 3)]
        say(f"  {label} vs fp64: dx {e[0]:.2e} dv {e[1]:.2e} dm {e[2]:.2e}")
        assert max(e) <= 1e-5
        assert res[label][0] == pytest.approx(res["fp64"][0], rel=1e-9)
    e = [relnorm(res["sym"][q], res["onepass"][q]) for q in (1, 2, 3)]
    say(f"  sym vs onepass: dx {e[0]:.2e} dv {e[1]:.2e} dm {e[2]:.2e}")
    assert max(e) <= 1e-5


@pytest.mark.parametrize("pattern", ["equal", "one_heavy", "few_heavy"])
def test_symmetric_vjp_mass_patterns(pattern):
    """Blocks whose targets and sources share one mass take the symmetric
    VJP's mass-free pair form; a few differing masses leave the blocks that
    hold them on the general form.  Both against the FP64 adjoint."""
    n, steps, eps2, dt = 5000, 6, 1e-4, 1e-3
    pos, vel, tgt = small_case(n, steps, seed=11)
    pos = mass_pattern(pos, pattern)
    res = {}
    for label, prec, tile in (("fp64", U.PREC_FP64, 0), ("sym", U.PREC_MIXED, tile_id("sym4r32"))):
        say(f"case: VJP {label} masses={pattern} n={n}")
        with U.System(n, prec, eps2, dt, tile=tile) as s:
            s.set_state(pos, vel)
            res[label] = s.adjoint(steps, 3, tgt)
    e = [relnorm(res["sym"][q], res["fp64"][q]) for q in (1, 2, 3)]
    say(f"  sym vs fp64: dx {e[0]:.2e} dv {e[1]:.2e} dm {e[2]:.2e}")
    assert max(e) <= 1e-5
    assert res["sym"][0] == pytest.approx(res["fp64"][0], rel=1e-9)


def test_autograd_function_wrapper():
    """A torch.autograd.Function whose forward and backward run through the C
    ABI (utcnb_adjoint); it lives only in this test harness."""
    import torch

    class NBodyLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, v0, m, target, eps2, dt, steps, ckpt):
            n = x0.shape[0]
            pos = np.concatenate([x0.detach().cpu().numpy(), m.detach().cpu().numpy()[:, None]], 1)
            vel = np.concatenate([v0.detach().cpu().numpy(), np.ones((n, 1))], 1)
            with U.System(n, U.PREC_MIXED, eps2, dt) as s:
                s.set_state(pos, vel)
                L, dx, dv, dm = s.adjoint(steps, ckpt, target.detach().cpu().numpy())
            ctx.save_for_backward(torch.from_numpy(dx), torch.from_numpy(dv), torch.from_numpy(dm))
            return torch.tensor(L, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gL):
            dx, dv, dm = ctx.saved_tensors
            return gL * dx, gL * dv, gL * dm, None, None, None, None, None

    n, steps, eps2, dt = 512, 10, 1e-3, 1e-3
    pos, vel, tgt = small_case(n, steps)
    x0 = torch.tensor(pos[:, :3], requires_grad=True)
    v0 = torch.tensor(vel[:, :3], requires_grad=True)
    m = torch.tensor(pos[:, 3], requires_grad=True)
    say("case: autograd wrapper")
    L = NBodyLoss.apply(x0, v0, m, torch.tensor(tgt), eps2, dt, steps, 5)
    (2.0 * L).backward()
    L0, dx0, dv0, dm0 = O.adjoint_f64(pos, vel, tgt, steps, eps2, dt, 1.0, 1)
    assert float(L) == pytest.approx(L0, rel=1e-9)
    assert relnorm(x0.grad.numpy(), 2 * dx0) <= 1e-5
    assert relnorm(v0.grad.numpy(), 2 * dv0) <= 1e-5
    assert relnorm(m.grad.numpy(), 2 * dm0) <= 1e-5
