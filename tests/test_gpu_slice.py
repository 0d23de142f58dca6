"""GPU tests of the target-slice path: one GPU's share of a system sharded over
targets, as the reference's multi-GPU nbody runs it
(benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.cu:41-49 partition,
bodysystem_kernel.cu:88-140 slice kernel).

* the kernel-level slice step (utcnb_integrate_slice_f32/f64, slice-local
  pos_new and vel) is bit-identical to the oracle's rows of the reference's
  force loop and integrator, on uneven slices of N = 262147;
* slices covering [0, n) and a commit equal one full step with the same
  one-pass tile, bit for bit, in every precision;
* four handles on one GPU, each stepping its own slice and receiving the
  others' slices into its next position buffer (the all-gather of a sharded
  run), end bitwise where one handle's full steps end;
* the slice field against the CPU FP64 oracle's rows;
* the shares of a symmetric-tile system (wave blocks, each unordered pair
  once, reactions summed over the shares) follow the full symmetric step to
  1e-12, and four handles running one share each with the reduce-scatter and
  all-gather done between them end bitwise where one handle running every
  share ends.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import utc_amd as U
from nbcases import plummer, relnorm
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_partition_rule():
    assert U.partition(262147, 8)[0] == (0, 32769)
    assert U.partition(262147, 8)[3] == (3 * 32769, 32768)
    assert U.partition(262147, 8)[7] == (262147 - 32768, 32768)
    parts = U.partition(1001, 7)
    assert sum(c for _, c in parts) == 1001
    assert all(parts[k][0] + parts[k][1] == parts[k + 1][0] for k in range(6))


@pytest.mark.parametrize("which", [0, 7])
def test_kernel_level_slice_bit_exact_f32(which):
    """N = 262147 over 8 shares (the first three take 32769 bodies, the rest
    32768): share 0 and share 7 in the reference's FP32 arithmetic against the
    oracle's rows, bit for bit."""
    import torch
    n = 262147
    start, cnt = U.partition(n, 8)[which]
    c = plummer(n, eps=0.002, dt=2.5e-4)
    pos = c["pos"].astype(np.float32)
    vel = c["vel"].astype(np.float32)
    eps2 = np.float32(0.002) * np.float32(0.002)
    dt, damping = np.float32(2.5e-4), np.float32(0.995)
    say(f"case: kernel-level slice f32 n={n} share={which} start={start} count={cnt}")
    d_old = torch.from_numpy(pos).cuda()
    d_new = torch.zeros((cnt, 4), dtype=torch.float32, device="cuda")
    d_vel = torch.from_numpy(vel[start:start + cnt].copy()).cuda()
    torch.cuda.synchronize()
    U.check(U.lib().utcnb_integrate_slice_f32(d_old.data_ptr(), d_new.data_ptr(), d_vel.data_ptr(), n, start,
                                              cnt, dt, eps2, damping, 0, None), "utcnb_integrate_slice_f32")
    torch.cuda.synchronize()
    f = O.gravitation_rows_f32(pos, eps2, start, start + cnt, nthreads=16)[start:start + cnt].copy()
    op, ov = pos[start:start + cnt].copy(), vel[start:start + cnt].copy()
    O.lib().orc_integrate_f32(cnt, O._ptr(op), O._ptr(ov), O._ptr(f), C.c_float(dt), C.c_float(damping))
    assert np.array_equal(d_new.cpu().numpy(), op)
    assert np.array_equal(d_vel.cpu().numpy(), ov)


def test_kernel_level_slice_bit_exact_f64():
    import torch
    n = 5003
    c = plummer(n, eps=0.01, dt=1e-3)
    pos, vel = c["pos"], c["vel"]
    eps2, dt, damping = 1e-4, 1e-3, 0.995
    f = O.gravitation_f64(pos, eps2)
    d_old = torch.from_numpy(pos).cuda()
    for start, cnt in U.partition(n, 3):
        say(f"case: kernel-level slice f64 n={n} start={start} count={cnt}")
        d_new = torch.zeros((cnt, 4), dtype=torch.float64, device="cuda")
        d_vel = torch.from_numpy(vel[start:start + cnt].copy()).cuda()
        torch.cuda.synchronize()
        U.check(U.lib().utcnb_integrate_slice_f64(d_old.data_ptr(), d_new.data_ptr(), d_vel.data_ptr(), n,
                                                  start, cnt, dt, eps2, damping, 128, None),
                "utcnb_integrate_slice_f64")
        torch.cuda.synchronize()
        op, ov = pos[start:start + cnt].copy(), vel[start:start + cnt].copy()
        fs = f[start:start + cnt].copy()
        O.lib().orc_integrate_f64(cnt, O._ptr(op), O._ptr(ov), O._ptr(fs), C.c_double(dt), C.c_double(damping))
        assert np.array_equal(d_new.cpu().numpy(), op)
        assert np.array_equal(d_vel.cpu().numpy(), ov)


def test_reference_accel_slice_bit_exact():
    """utcnb_accel_slice of a REFERENCE handle (FP32 and FP64 state): the
    reference's field rows, bit for bit the oracle's."""
    n = 5003
    c = plummer(n, eps=0.01, dt=1e-3)
    start, cnt = U.partition(n, 4)[2]
    say(f"case: reference accel_slice n={n} start={start} count={cnt}")
    pos32 = c["pos"].astype(np.float32)
    eps2f = np.float32(0.01) * np.float32(0.01)
    with U.System(n, U.PREC_REFERENCE, float(eps2f), 1e-3, kick=U.KICK_REFERENCE, state_dtype=U.F32) as s:
        s.set_state(pos32)
        got = s.accel_slice(start, cnt)
        ptr, dt = s.position_buffer(0)
        assert ptr and dt == U.F32
    want = O.gravitation_rows_f32(pos32, eps2f, start, start + cnt, nthreads=8)[start:start + cnt]
    assert np.array_equal(got, want.astype(np.float64))
    with U.System(n, U.PREC_REFERENCE, 1e-4, 1e-3, kick=U.KICK_REFERENCE, state_dtype=U.F64) as s:
        s.set_state(c["pos"])
        got = s.accel_slice(start, cnt)
        assert s.position_buffer(1)[1] == U.F64
    want = O.gravitation_f64(c["pos"], 1e-4)[start:start + cnt]
    assert np.array_equal(got, want)


def test_kernel_level_slice_rejects_bad_ranges():
    L = U.lib()
    for start, cnt in ((-1, 4), (0, 0), (8, 3), (5, 6)):
        assert L.utcnb_integrate_slice_f32(1, 2, 3, 10, start, cnt, 0.1, 0.1, 1.0, 0, None) == U.EINVAL
    assert b"start" in L.utcnb_last_error()


CASES = [
    ("mixed", U.PREC_MIXED, "k1s16", U.F64),
    ("fp32", U.PREC_FP32, "k2s8", U.F64),
    ("small", U.PREC_MIXED, "sm1t16w16n8k", U.F64),
    ("small_fp32", U.PREC_FP32, "sm2t16w16n8k", U.F64),
    ("fp64", U.PREC_FP64, None, U.F64),
    ("ref_f32", U.PREC_REFERENCE, None, U.F32),
    ("ref_f64", U.PREC_REFERENCE, None, U.F64),
]


@pytest.mark.parametrize("label,prec,tile,sdt", CASES)
def test_slices_then_commit_equal_full_steps(label, prec, tile, sdt):
    n, steps, shares = 5003, 3, 5
    c = plummer(n, eps=0.01, dt=1e-3)
    kick = U.KICK_REFERENCE if prec == U.PREC_REFERENCE else U.KICK_PHYSICAL
    tid = 0 if tile is None else U.tiles().index(tile) + 1
    pos = c["pos"].astype(np.float32) if sdt == U.F32 else c["pos"]
    vel = c["vel"].astype(np.float32) if sdt == U.F32 else c["vel"]
    dtype = np.float32 if sdt == U.F32 else np.float64
    say(f"case: slices+commit vs full steps {label} n={n}")
    out = []
    for sliced in (False, True):
        with U.System(n, prec, 1e-4, 1e-3, 0.999, kick=kick, state_dtype=sdt, tile=tid) as s:
            s.set_state(pos, vel)
            if tile is not None:
                assert s.slice_tile(100)[1] == tile
            for _ in range(steps):
                if sliced:
                    for start, cnt in U.partition(n, shares):
                        s.step_slice(start, cnt)
                    s.commit_step()
                else:
                    s.step(1)
            out.append(s.get_state(dtype))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipDeviceSynchronize.restype = C.c_int
    return h


def test_sharded_handles_with_all_gather_equal_one_system():
    """Four MIXED handles on one GPU stand for four GPUs: each steps its own
    share, then receives every other share's new FP32 positions into its next
    position buffer (what an all-gather over the GPUs would do) and commits.
    After 4 steps the shares' FP64 states, put together, are bitwise the state
    of one handle running full steps with the same one-pass tile."""
    n, G, steps = 20001, 4, 4
    c = plummer(n, eps=0.005, dt=5e-4)
    tid = U.tiles().index("k1s16") + 1
    parts = U.partition(n, G)
    hip = _hip()
    say(f"case: {G} sharded handles n={n}")
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], tile=tid) as one:
        one.set_state(c["pos"], c["vel"])
        one.step(steps)
        ref_p, ref_v = one.get_state()
    hs = [U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], tile=tid) for _ in range(G)]
    try:
        for h in hs:
            h.set_state(c["pos"], c["vel"])
        for _ in range(steps):
            for h, (start, cnt) in zip(hs, parts):
                h.step_slice(start, cnt)
            for h in hs:
                h.sync()
            for g, h in enumerate(hs):  # the all-gather of the new snapshots
                dst, dt = h.position_buffer(1)
                assert dt == U.F32
                for r, (start, cnt) in enumerate(parts):
                    if r == g:
                        continue
                    src, _ = hs[r].position_buffer(1)
                    assert hip.hipMemcpy(dst + 16 * start, src + 16 * start, 16 * cnt, 3) == 0
            assert hip.hipDeviceSynchronize() == 0
            for h in hs:
                h.commit_step()
        got_p = np.zeros_like(ref_p)
        got_v = np.zeros_like(ref_v)
        for h, (start, cnt) in zip(hs, parts):
            p, v = h.get_state()
            got_p[start:start + cnt] = p[start:start + cnt]
            got_v[start:start + cnt] = v[start:start + cnt]
    finally:
        for h in hs:
            h.close()
    assert np.array_equal(got_p, ref_p)
    assert np.array_equal(got_v, ref_v)


@pytest.mark.parametrize("which", [0, 7])
def test_slice_field_against_cpu_fp64_oracle(which):
    """The field of share 0 / share 7 of N = 262147 (FP64, mixed and plain
    FP32 on the slice's one-pass tile) against the CPU FP64 oracle's rows
    (512 sampled); the FP64 and one-pass slice fields equal the rows of the
    full evaluation bit for bit."""
    n = 262147
    start, cnt = U.partition(n, 8)[which]
    c = plummer(n, eps=0.002, dt=2.5e-4)
    eps2 = c["eps"] ** 2
    rows = np.sort(np.random.default_rng(which).choice(cnt, 512, replace=False))
    ref = O.accel_rows_f64(c["pos"], eps2, start + rows, nthreads=16)
    say(f"case: slice field vs CPU FP64 n={n} share={which}")
    errs = {}
    for label, prec in (("fp64", U.PREC_FP64), ("mixed", U.PREC_MIXED), ("fp32", U.PREC_FP32)):
        with U.System(n, prec, eps2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
            s.set_state(c["pos"], c["vel"])
            a = s.accel_slice(start, cnt)
            errs[label] = relnorm(a[rows], ref)
            if prec == U.PREC_FP64:
                assert np.array_equal(a, s.accel()[start:start + cnt])
        if prec != U.PREC_FP64:
            tid = U.tiles().index("k1s16") + 1
            with U.System(n, prec, eps2, c["dt"], flags=U.FLAG_NO_NEWTON, tile=tid) as s:
                s.set_state(c["pos"], c["vel"])
                assert np.array_equal(s.accel_slice(start, cnt), s.accel()[start:start + cnt])
    say(f"  rel err: {errs}")
    assert errs["fp64"] <= 1e-14
    assert errs["mixed"] <= 2e-6
    assert errs["fp32"] <= 4e-6


# ---- shares of a symmetric-tile system (each unordered pair once) -----------------
def _share_step_all(s, G, n):
    """One step of every share of one handle (the data flow of a G-GPU run
    with the reduce-scatter and the all-gather done in place)."""
    import torch
    reacts = [torch.empty((n, 3), dtype=torch.float64, device="cuda") for _ in range(G)]
    torch.cuda.synchronize()
    for g in range(G):
        s.share_pairs(G, g, reacts[g].data_ptr())
    s.sync()
    total = reacts[0].clone()
    for g in range(1, G):
        total += reacts[g]
    torch.cuda.synchronize()
    sums = []
    for g in range(G):
        start, cnt = s.share_range(G, g)
        sums.append(total[start:start + cnt].contiguous())
    torch.cuda.synchronize()
    for g in range(G):
        s.share_finish(G, g, sums[g].data_ptr())
    s.sync()
    s.commit_step()


@pytest.mark.parametrize("n,G", [(65536, 8), (20001, 3)])
def test_shares_equal_full_symmetric_steps(n, G):
    """Every share of one handle, their reactions summed, then every share
    finished and a commit: the full step's partial sums in another FP64 order,
    so the trajectory matches utcnb_step with the same tile to 1e-12, and (at
    the smaller N, where the CPU is quick) the FP64 CPU oracle to 1e-9 after
    3 steps."""
    c = plummer(n, eps=0.005, dt=5e-4)
    steps = 3
    say(f"case: shares vs full symmetric steps n={n} G={G}")
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as full:
        assert full.tile()[1].startswith("sym")
        full.set_state(c["pos"], c["vel"])
        full.step(steps)
        pf, vf = full.get_state()
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
        s.set_state(c["pos"], c["vel"])
        ranges = [s.share_range(G, g) for g in range(G)]
        assert ranges[0][0] == 0 and sum(cnt for _, cnt in ranges) == n
        assert all(ranges[g][0] + ranges[g][1] == ranges[g + 1][0] for g in range(G - 1))
        for _ in range(steps):
            _share_step_all(s, G, n)
        ps, vs = s.get_state()
    err = relnorm(ps[:, :3], pf[:, :3])
    say(f"  shares vs full step: position rel {err:.3e}")
    assert err <= 1e-12
    assert relnorm(vs[:, :3], vf[:, :3]) <= 1e-10
    if n <= 32768:
        p64, _ = O.steps_f64(c["pos"], c["vel"], steps, c["eps"] ** 2, c["dt"])
        assert relnorm(ps[:, :3], p64[:, :3]) <= 1e-9


def test_sharded_symmetric_handles_equal_one_handle():
    """Four handles on one GPU stand for four GPUs: each runs its own share,
    the shares' reaction arrays are summed (the reduce-scatter), each finishes
    its share, and the new FP32 positions are gathered into every handle's
    next buffer (the all-gather).  Bitwise the state of one handle running
    every share itself."""
    import torch
    n, G, steps = 20001, 4, 3
    c = plummer(n, eps=0.005, dt=5e-4)
    hip = _hip()
    say(f"case: {G} sharded symmetric handles n={n}")
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as one:
        one.set_state(c["pos"], c["vel"])
        for _ in range(steps):
            _share_step_all(one, G, n)
        ref_p, ref_v = one.get_state()
    hs = [U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) for _ in range(G)]
    try:
        for h in hs:
            h.set_state(c["pos"], c["vel"])
        ranges = [hs[0].share_range(G, g) for g in range(G)]
        reacts = [torch.empty((n, 3), dtype=torch.float64, device="cuda") for _ in range(G)]
        for _ in range(steps):
            torch.cuda.synchronize()
            for g, h in enumerate(hs):
                h.share_pairs(G, g, reacts[g].data_ptr())
            for h in hs:
                h.sync()
            total = reacts[0].clone()
            for g in range(1, G):
                total += reacts[g]
            sums = [total[a:a + k].contiguous() for a, k in ranges]
            torch.cuda.synchronize()
            for g, h in enumerate(hs):
                h.share_finish(G, g, sums[g].data_ptr())
            for h in hs:
                h.sync()
            for g, h in enumerate(hs):
                dst, _ = h.position_buffer(1)
                for r, (a, k) in enumerate(ranges):
                    if r != g:
                        src, _ = hs[r].position_buffer(1)
                        assert hip.hipMemcpy(dst + 16 * a, src + 16 * a, 16 * k, 3) == 0
            assert hip.hipDeviceSynchronize() == 0
            for h in hs:
                h.commit_step()
        got_p = np.zeros_like(ref_p)
        got_v = np.zeros_like(ref_v)
        for h, (a, k) in zip(hs, ranges):
            p, v = h.get_state()
            got_p[a:a + k] = p[a:a + k]
            got_v[a:a + k] = v[a:a + k]
    finally:
        for h in hs:
            h.close()
    assert np.array_equal(got_p, ref_p)
    assert np.array_equal(got_v, ref_v)


def test_share_calls_reject_other_systems():
    L = U.lib()
    tid = U.tiles().index("k2s8") + 1
    a, b = C.c_int(), C.c_int()
    with U.System(4096, U.PREC_MIXED, 1e-4, 1e-3, tile=tid) as s:
        assert L.utcnb_share_range(s.h, 2, 0, C.byref(a), C.byref(b)) == U.EINVAL
        assert b"symmetric" in L.utcnb_last_error()
    with U.System(4096, U.PREC_MIXED, 1e-4, 1e-3, tile=U.tiles().index("sym2r32") + 1) as s:
        assert L.utcnb_share_range(s.h, 0, 0, C.byref(a), C.byref(b)) == U.EINVAL
        assert L.utcnb_share_range(s.h, 2, 2, C.byref(a), C.byref(b)) == U.EINVAL
        assert L.utcnb_share_range(s.h, 100000, 0, C.byref(a), C.byref(b)) == U.EINVAL


# ---- the symmetric step in chunks of wave blocks (less partial-sum memory) ---------
def test_chunked_symmetric_step_equals_two_shares():
    """UTCNB_FLAG_SYM_CHUNKS at N = 20001: the step runs as two chunks of wave
    blocks, one after the other, the reactions accumulated in FP64; bitwise the
    two-share flow on one handle, and within 1e-12 of the single-launch
    symmetric step; the field within 1e-7 of the single launch's."""
    n, steps = 20001, 3
    c = plummer(n, eps=0.005, dt=5e-4)
    say(f"case: chunked symmetric step n={n}")
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON | U.FLAG_SYM_CHUNKS) as ch:
        assert ch.sym_chunks() == 2
        ch.set_state(c["pos"], c["vel"])
        a_ch = ch.accel()
        ch.step(steps)
        pc, vc = ch.get_state()
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as sh:
        assert sh.sym_chunks() == 0
        sh.set_state(c["pos"], c["vel"])
        a_full = sh.accel()
        for _ in range(steps):
            _share_step_all(sh, 2, n)
        ps, vs = sh.get_state()
    assert np.array_equal(pc, ps)
    assert np.array_equal(vc, vs)
    assert relnorm(a_ch, a_full) <= 1e-7
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as full:
        full.set_state(c["pos"], c["vel"])
        full.step(steps)
        pf, _ = full.get_state()
    assert relnorm(pc[:, :3], pf[:, :3]) <= 1e-12


def test_large_n_takes_chunked_symmetric_steps():
    """N = 2.5M: the whole step's reaction partials (73 GB) exceed the device
    budget, so the default MIXED system keeps the symmetric tile and evaluates
    it in chunks; its field at 512 sampled bodies against the CPU FP64 oracle,
    and one step against the field's kick-drift."""
    n = 2_500_000
    c = plummer(n, eps=0.002, dt=2.5e-4)
    rows = np.sort(np.random.default_rng(5).choice(n, 512, replace=False))
    say(f"case: large N={n} chunked symmetric default")
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
        name, G = s.tile()[1], s.sym_chunks()
        say(f"  tile {name}, {G} chunks")
        assert name.startswith("sym") and G >= 2
        s.set_state(c["pos"], c["vel"])
        a = s.accel()
        s.step(1)
        p1, v1 = s.get_state()
    ref = O.accel_rows_f64(c["pos"], c["eps"] ** 2, rows, nthreads=16)
    err = relnorm(a[rows], ref)
    say(f"  field rel err vs CPU FP64 oracle: {err:.3e}")
    assert err <= 2e-6
    # physical kick: v1 = v0 + a dt, x1 = x0 + v1 dt (FP64 state, the field above)
    v_want = c["vel"][:, :3] + a * c["dt"]
    assert relnorm(v1[:, :3], v_want) <= 1e-12
    assert np.all(np.isfinite(p1))


# ---- edge cases --------------------------------------------------------------------
def test_slice_edges_single_bodies():
    """Slices of one body (the first, the last) and a one-body system, in the
    reference arithmetic and MIXED: each equals the full step's body."""
    n = 777
    c = plummer(n, eps=0.01, dt=1e-3)
    say("case: one-body slices")
    for prec in (U.PREC_REFERENCE, U.PREC_MIXED):
        kick = U.KICK_REFERENCE if prec == U.PREC_REFERENCE else U.KICK_PHYSICAL
        tid = 0 if prec == U.PREC_REFERENCE else U.tiles().index("k1s16") + 1
        with U.System(n, prec, 1e-4, 1e-3, kick=kick, tile=tid) as full:
            full.set_state(c["pos"], c["vel"])
            full.step(1)
            pf, vf = full.get_state()
        with U.System(n, prec, 1e-4, 1e-3, kick=kick, tile=tid) as s:
            s.set_state(c["pos"], c["vel"])
            s.step_slice(0, 1)
            s.step_slice(n - 1, 1)
            s.commit_step()
            p, v = s.get_state()
        for i in (0, n - 1):
            assert np.array_equal(p[i], pf[i]) and np.array_equal(v[i], vf[i]), (prec, i)
    with U.System(1, U.PREC_MIXED, 1e-4, 1e-3) as s:
        s.set_state(c["pos"][:1], c["vel"][:1])
        s.step_slice(0, 1)
        s.commit_step()
        p, _ = s.get_state()
        a = s.accel_slice(0, 1)
    assert np.array_equal(a, np.zeros((1, 3)))  # a lone body feels no force
    assert np.allclose(p[0, :3], c["pos"][0, :3] + c["vel"][0, :3] * 1e-3, rtol=0, atol=1e-15)


def test_shares_with_padding_only_shares():
    """N = 1025 on sym2r32 padded to 2048: 16 wave blocks of 128 bodies, 7 of
    them padding only; 16 shares, the last 7 with no bodies: the all-shares
    flow still follows the single-launch step (1e-12), and the chunked step
    (two chunks) as well."""
    n, G = 1025, 16
    c = plummer(n, eps=0.01, dt=1e-3)
    tid = U.tiles().index("sym2r32") + 1
    say(f"case: shares with padding-only shares n={n} G={G}")
    with U.System(n, U.PREC_MIXED, 1e-4, 1e-3, tile=tid) as full:
        full.set_state(c["pos"], c["vel"])
        full.step(2)
        pf, _ = full.get_state()
    with U.System(n, U.PREC_MIXED, 1e-4, 1e-3, tile=tid) as s:
        ranges = [s.share_range(G, g) for g in range(G)]
        assert [k for _, k in ranges[9:]] == [0] * 7 and ranges[8] == (1024, 1)
        s.set_state(c["pos"], c["vel"])
        for _ in range(2):
            _share_step_all(s, G, n)
        ps, _ = s.get_state()
    assert relnorm(ps[:, :3], pf[:, :3]) <= 1e-12
    with U.System(n, U.PREC_MIXED, 1e-4, 1e-3, tile=tid, flags=U.FLAG_SYM_CHUNKS) as s:
        assert s.sym_chunks() == 2
        s.set_state(c["pos"], c["vel"])
        s.step(2)
        pc, _ = s.get_state()
    assert relnorm(pc[:, :3], pf[:, :3]) <= 1e-12
