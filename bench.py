#!/usr/bin/env python3
"""Benchmark of the nbody hot path on one MI355X.

Workload (default): N = 262144 bodies, Plummer sphere (a = 1, M = 1, equal
masses, seed 0), eps = 0.002, dt = 2.5e-4, physical kick, K timed steps.  A
step is one all-pairs force evaluation with the integrator fused into its last
kernel (at this size: the symmetric pair kernel, then the kernel that sums its
partials and applies the FP64 kick-drift); the timed region covers exactly K
steps with the state resident in HBM.

Prints ONE JSON line (stdout).  `value` is interactions/s of the mixed path
(FP32 pairs with compensated sums, FP64 state): N^2 * K / wall time.  Also
reported: the plain FP32 path, the Newton-step variant, accuracy against the
on-device FP64 path (acceleration and final-position relative error), energy
drift, forward+adjoint throughput, the roofline of the force kernel and a
CPU baseline (the oracle's restatement of the reference's sequential CPU path,
threaded over bodies with the reference's partition rule, on a row sample).

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 5] [--n 262144]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "interactions/s fwd & fwd+adjoint at N=262144 vs plain FP32 GPU; rel err"
PEAK_FP32 = 157.3e12      # MI355X FP32 vector peak (AMD spec), FLOP/s
FLOP_PER_PAIR = 20        # reference convention, benchapps/nbody/seq/bodysystem.h:75-92
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def relnorm(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def device_sync(s):
    """The system's stream, then the whole device (torch.cuda.synchronize)."""
    import torch
    s.sync()
    torch.cuda.synchronize()


def timed_steps(U, prec, n, c, steps, warmup, flags=0, tile=0):
    """Run warmup + timed steps; returns (wall_ms, kernel_ms, launches, system)."""
    s = U.System(n, prec, c["eps"] ** 2, c["dt"], flags=flags, tile=tile)
    s.set_state(c["pos"], c["vel"])
    if warmup:
        s.step(warmup)
    s.sync()
    # reset to the initial state so the timed steps are the configuration's steps
    s.set_state(c["pos"], c["vel"])
    device_sync(s)
    t0 = time.perf_counter()
    s.step(steps)
    device_sync(s)
    wall = (time.perf_counter() - t0) * 1e3
    kms, launches = s.last_kernel_ms()
    return wall, kms, launches, s


def read_traffic(path: Path):
    """HBM bytes per launch of the force kernel from the committed PMC summary."""
    if not path.exists():
        return None
    txt = path.read_text()
    m = re.search(r"traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def plummer65536_accuracy(U):
    """N = 65536 Plummer, eps = 0.005, dt = 5e-4, 20 steps: plain FP32 and mixed
    against the on-device FP64 path (acceleration and final positions)."""
    from tests.nbcases import plummer65536
    c = plummer65536()
    n, eps2, steps = len(c["pos"]), c["eps"] ** 2, c["steps"]
    with U.System(n, U.PREC_FP64, eps2, c["dt"]) as s64:
        s64.set_state(c["pos"], c["vel"])
        a64 = s64.accel()
        s64.step(steps)
        p64, _ = s64.get_state()
    out = {"workload": f"N={n} Plummer, eps=0.005, dt=5e-4, {steps} steps, physical kick"}
    for label, prec in (("mixed", U.PREC_MIXED), ("fp32", U.PREC_FP32)):
        with U.System(n, prec, eps2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
            s.set_state(c["pos"], c["vel"])
            a = s.accel()
            s.set_state(c["pos"], c["vel"])
            s.sync()
            t0 = time.perf_counter()
            s.step(steps)
            s.sync()
            wall = (time.perf_counter() - t0) * 1e3
            p, _ = s.get_state()
            out[label] = dict(tile=s.tile()[1], ms_per_step=wall / steps,
                              interactions_per_s=float(n) * n * steps / (wall * 1e-3),
                              accel_relerr=relnorm(a, a64), final_pos_relerr=relnorm(p[:, :3], p64[:, :3]))
        log(f"N=65536 Plummer {label}: {out[label]['ms_per_step']:.3f} ms/step, accel rel err "
            f"{out[label]['accel_relerr']:.3e}, final-position rel err {out[label]['final_pos_relerr']:.3e}")
    return out


def cpu_baseline(c, n, threads, rows):
    """The oracle's restatement of benchapps/nbody/seq/bodysystem.cc:103-133
    (bit-identical to the reference), rows [0, rows) x all n sources, one
    force evaluation, threaded over rows with the reference's partition rule."""
    from oracle import oracle as O
    pos32 = c["pos"].astype(np.float32)
    eps2 = np.float32(c["eps"]) * np.float32(c["eps"])
    O.gravitation_rows_f32(pos32, eps2, 0, min(256, rows), threads)  # warm the pages
    t0 = time.perf_counter()
    O.gravitation_rows_f32(pos32, eps2, 0, rows, threads)
    dt = time.perf_counter() - t0
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = {
        "value": rows * float(n) / dt,
        "unit": "interactions/s",
        "cores": threads,
        "host_cpus": os.cpu_count(),
        "cpu_model": model,
        "threads_pinned": True,  # worker t on the t-th CPU of the process's affinity set
        "kind": "port",
        "sample": f"one force evaluation of rows 0..{rows - 1} against all {n} sources "
                  f"(FP32, reference order) = {rows * float(n):.3g} pair interactions, "
                  f"{dt:.2f} s wall on {threads} threads",
    }
    # the reference's own sequential code (benchapps/nbody/seq/bodysystem.cc,
    # compiled from its sources into oracle/_ref), single-threaded as written
    if O.REF_LIB.exists():
        import ctypes as C
        m = 65536  # ~10 s single-threaded
        p = np.ascontiguousarray(pos32[:m])
        v = np.zeros((m, 4), np.float32)
        v[:, 3] = 1.0
        t0 = time.perf_counter()
        O.ref_lib().ref_run_f32(m, 1, p.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
                                float(eps2), 1.0, 2.5e-4)
        dt1 = time.perf_counter() - t0
        out["reference_single_thread"] = {
            "value": float(m) * m / dt1, "unit": "interactions/s", "cores": 1, "kind": "reference",
            "sample": f"one step of the reference's BodySystem<float> on the first {m} bodies "
                      f"({float(m) * m:.3g} pairs), {dt1:.2f} s"}
    return out


def batched_256_systems(U, nsys=256, steps=20, seed=0):
    """256 Plummer systems, n ~ round(exp(U[ln 256, ln 8192])), each with its
    own seed, eps = 0.01, dt = 1e-3, 20 steps; forward and forward+adjoint, one
    captured HIP graph of persistent kernels vs per-system launches."""
    from tests.nbcases import batch256_sizes
    sizes = batch256_sizes(nsys, seed)
    pos, vel = [], []
    for k, n in enumerate(sizes):
        p, v = U.generate(U.GEN_PLUMMER, n, 1000 + k)
        pos.append(p)
        vel.append(v)
    pos, vel = np.concatenate(pos), np.concatenate(vel)
    tgt = U.generate(U.GEN_TARGETS, len(pos), 77, 0.5)
    pairs = float(sum(n * n for n in sizes)) * steps
    out = {"systems": nsys, "bodies": int(sum(sizes)), "pairs_per_step": pairs / steps, "steps": steps}
    # graph legs: batch tile 0 (the autotuner picks among the one-pass and the
    # symmetric batched kernels; systems padded to 256); per-system legs on a
    # batch padded to 64 only, as the single-system kernels need no more
    with U.Batch(sizes, 0.01 ** 2, 1e-3, tile=0) as bg, U.Batch(sizes, 0.01 ** 2, 1e-3, tile=1) as bp:
        for label, b, mode, adj in (("graph_fwd", bg, 0, False), ("per_system_fwd", bp, 1, False),
                                    ("graph_fwd_adj", bg, 0, True), ("per_system_fwd_adj", bp, 1, True)):
            b.set_state(pos, vel)
            if adj:
                b.run(steps, mode, 10, tgt)  # warm-up: capture / workspace / tuning
            else:
                b.run(steps, mode)
            b.set_state(pos, vel)
            ms = b.run(steps, mode, 10, tgt)[0] if adj else b.run(steps, mode)
            out[label + "_ms"] = ms
            out[label + "_interactions_per_s"] = pairs / (ms * 1e-3)
            log(f"batched {label}: {ms:.2f} ms, {pairs / (ms * 1e-3):.4e} interactions/s")
        out["graph_kernel"] = bg.kernel_name()
    out["graph_speedup_fwd"] = out["per_system_fwd_ms"] / out["graph_fwd_ms"]
    out["graph_speedup_fwd_adj"] = out["per_system_fwd_adj_ms"] / out["graph_fwd_adj_ms"]
    return out


def reference_arithmetic(U, c, n, steps=3):
    """The reference's own arithmetic on the GPU (bit-identical to
    benchapps/nbody/seq/bodysystem.cc: sqrtf, FP64-rounded divide, sequential
    j order per body), FP32 state, the headline bodies and softening."""
    pos = c["pos"].astype(np.float32)
    vel = c["vel"].astype(np.float32)
    eps2 = np.float32(c["eps"]) * np.float32(c["eps"])
    with U.System(n, U.PREC_REFERENCE, float(eps2), float(np.float32(c["dt"])), 1.0,
                  kick=U.KICK_REFERENCE, state_dtype=U.F32) as s:
        s.set_state(pos, vel)
        s.step(1)  # warm-up
        s.set_state(pos, vel)
        device_sync(s)
        t0 = time.perf_counter()
        s.step(steps)
        device_sync(s)
        wall = (time.perf_counter() - t0) * 1e3
        kms, launches = s.last_kernel_ms()
    out = dict(steps=steps, ms_per_step=wall / steps, kernel_ms_per_step=kms / launches,
               interactions_per_s=float(n) * n * steps / (wall * 1e-3),
               kernel="ref_step_kernel<float, float4, true>: one thread per body, sources in index "
                      "order, the reference's operation order; bit-identical to its CPU path")
    log(f"reference arithmetic: {out['ms_per_step']:.2f} ms/step, {out['interactions_per_s']:.4e} interactions/s")
    return out


def cube_logm16384_fwd_adjoint(U):
    """N = 16384 uniform cube, log-uniform masses in [0.5, 2]/N, eps = 0.01,
    dt = 1e-3, 50 steps, checkpoint every 10: the forward alone and forward +
    adjoint (same loss as fwd_adjoint), mixed precision, default tile."""
    from tests.nbcases import cube_logm16384
    c = cube_logm16384()
    n, steps, ckpt = len(c["pos"]), c["steps"], c["ckpt"]
    pairs = float(n) * n
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
        s.set_state(c["pos"], c["vel"])
        s.step(steps)  # warm-up
        s.set_state(c["pos"], c["vel"])
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        fwd = (time.perf_counter() - t0) * 1e3
        s.set_state(c["pos"], c["vel"])
        s.adjoint(steps, ckpt, c["target"])  # warm-up (workspace)
        s.set_state(c["pos"], c["vel"])
        s.sync()
        t0 = time.perf_counter()
        s.adjoint(steps, ckpt, c["target"])
        adj = (time.perf_counter() - t0) * 1e3
        _, launches = s.last_kernel_ms()
        tile = s.tile()[1]
    out = dict(workload=f"N={n} cube, log-uniform masses, eps=0.01, dt=1e-3, {steps} steps, "
                        f"checkpoint every {ckpt}", tile=tile,
               fwd_ms=fwd, fwd_interactions_per_s=pairs * steps / (fwd * 1e-3),
               fwd_adj_ms=adj, forward_equivalent_interactions_per_s=pairs * steps / (adj * 1e-3),
               pair_evaluations=launches, ratio_vs_forward=fwd / adj,
               total_ratio_vs_forward=fwd / adj * launches / steps)
    log(f"N=16384 cube, log-uniform masses [{tile}]: forward {fwd:.2f} ms, fwd+adjoint {adj:.2f} ms "
        f"({out['ratio_vs_forward']:.3f}x forward-equivalent)")
    return out


def grad_fd_check(U, ncoords=64):
    """The gradient target: the mixed adjoint of the N = 16384 cube (log-uniform
    masses, 50 steps, checkpoint every 10) against FP64 central differences on
    64 random coordinates (positions, velocities, masses); each perturbed loss
    is an FP64 forward run on the GPU.  Steps as large as the loss's curvature
    allows: 1e-6 for a position, 1e-4 for a velocity, 5% of a mass."""
    from tests.nbcases import cube_logm16384
    c = cube_logm16384()
    n, steps, eps2 = len(c["pos"]), c["steps"], c["eps"] ** 2
    t0 = time.perf_counter()
    with U.System(n, U.PREC_MIXED, eps2, c["dt"], kick=c["kick"], flags=U.FLAG_NO_NEWTON) as s:
        s.set_state(c["pos"], c["vel"])
        _, dx, dv, dm = s.adjoint(steps, c["ckpt"], c["target"])
    rng = np.random.default_rng(11)
    kinds, bodies, comps = rng.integers(0, 3, ncoords), rng.integers(0, n, ncoords), rng.integers(0, 3, ncoords)
    errs = []

    def final(s, pos, vel):
        s.set_state(pos, vel)
        s.step(steps)
        return s.get_state()[0][:, :3]

    with U.System(n, U.PREC_FP64, eps2, c["dt"], kick=c["kick"]) as s:
        for kind, i, comp in zip(kinds, bodies, comps):
            comp = 3 if kind == 2 else comp
            arr = c["vel"] if kind == 1 else c["pos"]
            h = (1e-6, 1e-4, 0.05 * c["pos"][i, 3])[kind]
            a = arr.copy(); a[i, comp] += h
            b = arr.copy(); b[i, comp] -= h
            pp, pm = (final(s, c["pos"], a), final(s, c["pos"], b)) if kind == 1 else \
                     (final(s, a, c["vel"]), final(s, b, c["vel"]))
            fd = float(((pp - pm) * (pp + pm - 2.0 * c["target"])).sum() / (n * 2.0 * h))
            g = dm[i] if kind == 2 else (dv if kind == 1 else dx)[i, comp]
            errs.append(abs(g - fd) / abs(fd))
    out = {"workload": f"N={n} cube, log-uniform masses, {steps} steps, checkpoint every {c['ckpt']}, "
                       f"{ncoords} coordinates", "max_relerr": float(max(errs)),
           "median_relerr": float(np.median(errs)), "seconds": time.perf_counter() - t0}
    log(f"gradient vs FP64 central differences: max rel {out['max_relerr']:.3e} ({out['seconds']:.1f} s)")
    return out


def fwd_adjoint(U, c, n, K, fwd_rate, ckpt=10):
    """Forward K steps with checkpoints every `ckpt`, then the adjoint of the
    mean squared distance to seeded targets back to the initial state (mixed
    precision, default tile).  Rates: forward-equivalent N^2 K / wall, and all
    pair evaluations (forward, recompute, VJP, each N^2) / wall."""
    pairs = float(n) * float(n)
    tgt = U.generate(U.GEN_TARGETS, n, 0x5EED7A56E7, 0.5)
    sa = U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON)
    sa.set_state(c["pos"], c["vel"])
    sa.adjoint(K, ckpt, tgt)  # warm-up with the timed shape (allocates the workspace)
    sa.set_state(c["pos"], c["vel"])
    sa.sync()
    t0 = time.perf_counter()
    loss, dx, dv, dm = sa.adjoint(K, ckpt, tgt)
    wall = (time.perf_counter() - t0) * 1e3
    kms, launches = sa.last_kernel_ms()
    fe = pairs * K / (wall * 1e-3)
    out = dict(wall_ms=wall, ckpt_every=ckpt, loss=loss, tile=sa.tile()[1],
               forward_equivalent_interactions_per_s=fe,
               pair_evaluations=launches,
               total_pair_evaluations_per_s=pairs * launches / (wall * 1e-3))
    if fwd_rate:
        out["ratio_vs_forward"] = fe / fwd_rate
        out["total_ratio_vs_forward"] = pairs * launches / (wall * 1e-3) / fwd_rate
    log(f"fwd+adjoint (ckpt {ckpt}): {wall:.1f} ms, forward-equivalent {fe:.4e}/s "
        + (f"({fe / fwd_rate:.3f}x forward), " if fwd_rate else "")
        + f"total pair evaluations {launches}")
    return out


def slice_per_gpu(U, c, n, K, shares=8):
    """One GPU's share of the headline system sharded over `shares` GPUs by
    targets (the reference's multi-GPU scheme,
    benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.cu:41-49 and
    bodysystem_kernel.cu:88-140): share 0 (n/shares targets) against all n
    sources, mixed precision, one-pass tiles (the symmetric ones pair bodies of
    the whole system).  Every one-pass tile is timed; the default one and the
    fastest are reported."""
    start, cnt = U.partition(n, shares)[0]
    pairs = float(cnt) * float(n)
    names = U.tiles()
    sweep = {}
    for tid, name in enumerate(names, start=1):
        if not name.startswith("k"):  # one-pass tiles kKsS only (small-system tiles hold <= 8192 bodies)
            continue
        with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON, tile=tid) as s:
            s.set_state(c["pos"], c["vel"])
            for _ in range(10):  # warm-up
                s.step_slice(start, cnt)
            s.sync()
            t0 = time.perf_counter()
            for _ in range(K):
                s.step_slice(start, cnt)
                s.commit_step()
            s.sync()
            sweep[name] = (time.perf_counter() - t0) * 1e3 / K
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
        default = s.slice_tile(cnt)[1]
        s.set_state(c["pos"], c["vel"])
        s.step_slice(start, cnt)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(K):
            s.step_slice(start, cnt)
            s.commit_step()
        s.sync()
        ms = (time.perf_counter() - t0) * 1e3 / K
        kms, _ = s.last_kernel_ms()
    best = min(sweep, key=sweep.get)

    # the symmetric kernel's share of the same system: share 0 of `shares` by
    # wave blocks, each unordered pair once (utcnb_share_pairs, then
    # utcnb_share_finish with the reactions summed over all shares, computed
    # once before the timed loop; the reduce-scatter and all-gather of a
    # multi-GPU run are not part of one GPU's time)
    import torch
    sym = None
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
        s.set_state(c["pos"], c["vel"])
        tile = s.tile()[1]
        if tile.startswith("sym"):
            reacts = [torch.empty((n, 3), dtype=torch.float64, device="cuda") for _ in range(shares)]
            torch.cuda.synchronize()
            for g in range(shares):
                s.share_pairs(shares, g, reacts[g].data_ptr())
            s.sync()
            total = reacts[0].clone()
            for g in range(1, shares):
                total += reacts[g]
            s0, c0 = s.share_range(shares, 0)
            rsum = total[s0:s0 + c0].contiguous()
            torch.cuda.synchronize()
            # warm-up: about 50 ms of this loop (its first ~20 ms run ~10% slower
            # on MI355X, as the headline leg's first steps do before its warm-up)
            for _ in range(30):
                s.share_pairs(shares, 0, reacts[0].data_ptr())
                s.share_finish(shares, 0, rsum.data_ptr())
            s.sync()
            t0 = time.perf_counter()
            for _ in range(K):
                s.share_pairs(shares, 0, reacts[0].data_ptr())
                s.share_finish(shares, 0, rsum.data_ptr())
            s.sync()
            sms = (time.perf_counter() - t0) * 1e3 / K
            sp = float(c0) * float(n)
            sym = {"tile": tile, "shape": f"share 0 of {shares} by wave blocks: {c0} targets, each unordered pair "
                                           "once against the half ring of wave blocks after its own",
                   "targets": c0, "ms_per_step": sms, "interactions_per_s": sp / (sms * 1e-3),
                   "fraction_of_fp32_peak": FLOP_PER_PAIR * sp / (sms * 1e-3) / PEAK_FP32}
            log(f"symmetric share {c0}x{n} [{tile}]: {sms:.3f} ms/step, {sym['interactions_per_s']:.4e} "
                f"interactions/s, {sym['fraction_of_fp32_peak']:.3f} of peak")
    out = {
        "shape": f"{cnt} targets x {n} sources (share 0 of {shares})", "targets": cnt, "sources": n,
        "default_tile": default, "ms_per_step": ms, "kernel_ms_last_step": kms,
        "interactions_per_s": pairs / (ms * 1e-3),
        "fraction_of_fp32_peak": FLOP_PER_PAIR * pairs / (ms * 1e-3) / PEAK_FP32,
        "best_tile": best, "best_ms_per_step": sweep[best],
        "best_fraction_of_fp32_peak": FLOP_PER_PAIR * pairs / (sweep[best] * 1e-3) / PEAK_FP32,
        "tile_sweep_ms": sweep,
        "symmetric_share": sym,
    }
    log(f"slice {cnt}x{n} [{default}]: {ms:.3f} ms/step, {out['interactions_per_s']:.4e} interactions/s, "
        f"{out['fraction_of_fp32_peak']:.3f} of peak; best {best} {sweep[best]:.3f} ms")
    return out


def large_n(U, n=2_500_000, steps=2):
    """A system whose whole-step partial sums exceed the device budget (N =
    2.5M Plummer, eps = 0.002, dt = 2.5e-4, mixed): the default keeps the
    symmetric tile and evaluates each step in chunks of wave blocks."""
    from tests.nbcases import plummer
    c = plummer(n, eps=0.002, dt=2.5e-4)
    with U.System(n, U.PREC_MIXED, c["eps"] ** 2, c["dt"], flags=U.FLAG_NO_NEWTON) as s:
        s.set_state(c["pos"], c["vel"])
        s.step(1)  # warm-up
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        pairs = float(n) * float(n)
        out = {"n": n, "tile": s.tile()[1], "chunks": s.sym_chunks(), "ms_per_step": ms,
               "interactions_per_s": pairs / (ms * 1e-3),
               "fraction_of_fp32_peak": FLOP_PER_PAIR * pairs / (ms * 1e-3) / PEAK_FP32}
    log(f"N={n} [{out['tile']}, {out['chunks']} chunks]: {ms:.1f} ms/step, {out['interactions_per_s']:.4e} "
        f"interactions/s, {out['fraction_of_fp32_peak']:.3f} of peak")
    return out


def small_systems(U, sizes=(1024, 2048, 4096, 8192), steps=200):
    """Small single systems, where a step is bound by latency: uniform cube,
    equal masses 1/N, eps = 0.01, dt = 1e-3 (the N = 4096 case is the
    reference CPU path's workload), mixed and plain FP32, the default tile
    (one kernel per step) beside the symmetric sym2r32 (two kernels per step);
    time per step from the wall clock over `steps` steps after a warm-up."""
    sym = U.tiles().index("sym2r32") + 1
    out = []
    for n in sizes:
        pos, vel = U.generate(U.GEN_CUBE, n, 0, 0.01)
        for prec, label in ((U.PREC_MIXED, "mixed"), (U.PREC_FP32, "fp32")):
            row = {"n": n, "precision": label}
            for key, tile in (("default", 0), ("sym2r32", sym)):
                with U.System(n, prec, 1e-4, 1e-3, flags=U.FLAG_NO_NEWTON, tile=tile) as s:
                    s.set_state(pos, vel)
                    s.step(20)
                    s.sync()
                    t0 = time.perf_counter()
                    s.step(steps)
                    s.sync()
                    us = (time.perf_counter() - t0) * 1e6 / steps
                    row[key] = {"tile": s.tile()[1], "us_per_step": us,
                                "interactions_per_s": float(n) * n / (us * 1e-6)}
            row["speedup_vs_sym2r32"] = row["sym2r32"]["us_per_step"] / row["default"]["us_per_step"]
            out.append(row)
            log(f"N={n} {label}: {row['default']['tile']} {row['default']['us_per_step']:.2f} us/step, "
                f"sym2r32 {row['sym2r32']['us_per_step']:.2f} us/step")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=262144,
                    help="rows of the CPU baseline's force evaluation (default: all, ~10 s on 16 threads)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--only", choices=["mixed", "mixed_newton", "fp32", "fp32_newton", "adjoint", "batched", "reference",
                                       "slice", "large", "small"],
                    default=None,
                    help="time one path only and print its line (for profiler runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain 256 MiB device copy (to scale HBM counters) and exit")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        nbytes = 256 << 20
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 3}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is design-only in this build.", file=sys.stderr)
        sys.exit(2)

    import torch

    import utc_amd as U
    from tests.nbcases import plummer

    # bring up torch's device context before anything is timed (device_sync
    # uses it; its first use must not fall inside a timed region)
    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()

    n, K, W = args.n, args.steps, args.warmup
    c = plummer(n, eps=0.002, dt=2.5e-4, steps=K)
    name, cus, clk = U.device_info()
    log(f"device {name} CUs={cus} clock={clk} kHz; n={n} steps={K} warmup={W}")

    pairs = float(n) * float(n)
    if args.only == "batched":
        r = batched_256_systems(U)
        print(json.dumps({"metric": METRIC, "value": r["graph_fwd_interactions_per_s"],
                          "unit": "interactions/s", "n_gpus": 1, "only": "batched", "batched": r}),
              flush=True)
        return
    if args.only == "reference":
        r = reference_arithmetic(U, c, n, K)
        print(json.dumps({"metric": METRIC, "value": r["interactions_per_s"], "unit": "interactions/s",
                          "n_gpus": 1, "steps": K, "only": "reference", "reference_arithmetic": r}),
              flush=True)
        return
    if args.only == "large":
        r = large_n(U)
        print(json.dumps({"metric": METRIC, "value": r["interactions_per_s"], "unit": "interactions/s",
                          "n_gpus": 1, "only": "large", "large_n": r}), flush=True)
        return
    if args.only == "small":
        r = small_systems(U)
        print(json.dumps({"metric": METRIC, "value": r[0]["default"]["interactions_per_s"],
                          "unit": "interactions/s", "n_gpus": 1, "only": "small", "small_systems": r}),
              flush=True)
        return
    if args.only == "slice":
        r = slice_per_gpu(U, c, n, K)
        print(json.dumps({"metric": METRIC, "value": r["interactions_per_s"], "unit": "interactions/s",
                          "n_gpus": 1, "steps": K, "only": "slice", "slice_per_gpu": r}), flush=True)
        return
    if args.only == "adjoint":
        r = fwd_adjoint(U, c, n, K, None)
        print(json.dumps({"metric": METRIC, "value": r["forward_equivalent_interactions_per_s"],
                          "unit": "interactions/s", "n_gpus": 1, "steps": K, "warmup": W,
                          "only": "adjoint", "fwd_adjoint": r}), flush=True)
        return
    results = {}
    systems = {}
    variants = (("mixed", U.PREC_MIXED, U.FLAG_NO_NEWTON),
                ("mixed_newton", U.PREC_MIXED, 0),
                ("fp32", U.PREC_FP32, U.FLAG_NO_NEWTON),
                ("fp32_newton", U.PREC_FP32, 0))
    if args.only:
        variants = [v for v in variants if v[0] == args.only]
    for label, prec, flags in variants:
        wall, kms, launches, s = timed_steps(U, prec, n, c, K, W, flags=flags)
        results[label] = dict(wall_ms=wall, kernel_ms=kms, launches=launches, tile=s.tile()[1],
                              interactions_per_s=pairs * K / (wall * 1e-3),
                              kernel_interactions_per_s=pairs * launches / (kms * 1e-3))
        systems[label] = s
        log(f"{label} [{s.tile()[1]}]: {wall / K:.3f} ms/step wall, {kms / launches:.3f} ms/step kernel, "
            f"{results[label]['interactions_per_s']:.4e} interactions/s")

    # the same positions with log-uniform masses (0.1..10): no block has one
    # mass, so every pair takes the general form (the headline's equal masses,
    # as in the reference's initial conditions, take the mass-free form)
    unequal = None
    if not args.only:
        from tests.nbcases import mass_pattern
        cu = dict(c, pos=mass_pattern(c["pos"], "loguniform"))
        wall, kms, launches, su = timed_steps(U, U.PREC_MIXED, n, cu, K, W, flags=U.FLAG_NO_NEWTON)
        unequal = dict(masses="log-uniform in [0.1, 10], total 1", wall_ms=wall, kernel_ms=kms,
                       launches=launches, tile=su.tile()[1], interactions_per_s=pairs * K / (wall * 1e-3),
                       kernel_interactions_per_s=pairs * launches / (kms * 1e-3))
        su.close()
        log(f"mixed, unequal masses [{unequal['tile']}]: {wall / K:.3f} ms/step wall, "
            f"{unequal['interactions_per_s']:.4e} interactions/s")

    if args.only:
        r = results[args.only]
        print(json.dumps({"metric": METRIC, "value": r["interactions_per_s"], "unit": "interactions/s",
                          "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": r["wall_ms"] / K,
                          "only": args.only, "kernel_ms_per_step": r["kernel_ms"] / r["launches"]}),
              flush=True)
        return

    # ---- accuracy against the on-device FP64 path --------------------------------
    s64 = U.System(n, U.PREC_FP64, c["eps"] ** 2, c["dt"])
    s64.set_state(c["pos"], c["vel"])
    a64 = s64.accel()
    k0, w0, _ = s64.diagnostics()
    s64.step(K)
    p64, _ = s64.get_state()
    acc = {}
    for label, prec, flags in (("mixed", U.PREC_MIXED, U.FLAG_NO_NEWTON),
                               ("mixed_newton", U.PREC_MIXED, 0),
                               ("fp32", U.PREC_FP32, U.FLAG_NO_NEWTON),
                               ("fp32_newton", U.PREC_FP32, 0)):
        s = systems[label]
        p, _ = s.get_state()
        s.set_state(c["pos"], c["vel"])
        a = s.accel()
        acc[label] = dict(accel_relerr=relnorm(a, a64),
                          final_pos_relerr=relnorm(p[:, :3], p64[:, :3]),
                          final_pos_maxabs=float(np.abs(p[:, :3] - p64[:, :3]).max()))
        log(f"{label}: accel rel err {acc[label]['accel_relerr']:.3e}, final-position rel err "
            f"{acc[label]['final_pos_relerr']:.3e}")
    # energy drift of the mixed path over the K steps
    sm = systems["mixed"]
    sm.set_state(c["pos"], c["vel"])
    e0k, e0w, _ = sm.diagnostics()
    sm.step(K)
    e1k, e1w, p1 = sm.diagnostics()
    e0, e1 = e0k + e0w, e1k + e1w
    energy = dict(E0=e0, E1=e1, rel_drift=abs(e1 - e0) / abs(e0), virial_2K_over_W=2 * e0k / abs(e0w),
                  momentum_end=[float(x) for x in p1])
    log(f"energy drift {energy['rel_drift']:.3e}, virial 2K/|W| {energy['virial_2K_over_W']:.4f}")

    # ---- N = 65536 Plummer accuracy ------------------------------------------------
    cfg2 = None
    try:
        cfg2 = plummer65536_accuracy(U)
    except Exception as e:  # reported, never hidden
        cfg2 = {"error": str(e)}
        log(f"N=65536 Plummer accuracy failed: {e}")

    # ---- forward + adjoint ----------------------------------------------------------
    fwd_adj = None
    if not args.no_adjoint:
        try:
            fwd_adj = fwd_adjoint(U, c, n, K, results["mixed"]["interactions_per_s"])
            # a checkpoint every step: nothing to recompute (at K = 10 with checkpoints
            # every 10 the only segment is the last, whose positions the forward
            # sweep keeps, so the two legs do the same work)
            fwd_adj["store_every_step"] = fwd_adjoint(U, c, n, K, results["mixed"]["interactions_per_s"],
                                                      ckpt=1)
        except Exception as e:  # reported, never hidden
            fwd_adj = {"error": str(e)}
            log(f"fwd+adjoint failed: {e}")

    ref_arith = None
    try:
        ref_arith = reference_arithmetic(U, c, n)
    except Exception as e:  # reported, never hidden
        ref_arith = {"error": str(e)}
        log(f"reference arithmetic failed: {e}")

    cfg4 = None
    fdchk = None
    if not args.no_adjoint:
        try:
            cfg4 = cube_logm16384_fwd_adjoint(U)
        except Exception as e:  # reported, never hidden
            cfg4 = {"error": str(e)}
            log(f"N=16384 adjoint leg failed: {e}")
        try:
            fdchk = grad_fd_check(U)
        except Exception as e:  # reported, never hidden
            fdchk = {"error": str(e)}
            log(f"gradient check failed: {e}")

    # ---- one GPU's share of the system sharded over 8 GPUs ----------------------------
    slc = None
    try:
        slc = slice_per_gpu(U, c, n, K)
    except Exception as e:  # reported, never hidden
        slc = {"error": str(e)}
        log(f"slice leg failed: {e}")

    big = None
    try:
        big = large_n(U)
    except Exception as e:  # reported, never hidden
        big = {"error": str(e)}
        log(f"large-N leg failed: {e}")

    small = None
    try:
        small = small_systems(U)
    except Exception as e:  # reported, never hidden
        small = {"error": str(e)}
        log(f"small-systems leg failed: {e}")

    # ---- 256 batched systems -------------------------------------------------------
    batch = None
    if not args.no_batch:
        try:
            batch = batched_256_systems(U)
        except Exception as e:  # reported, never hidden
            batch = {"error": str(e)}
            log(f"batched run failed: {e}")

    # ---- roofline of the force kernel ------------------------------------------
    m = results["mixed"]
    t_launch = m["kernel_ms"] / m["launches"] * 1e-3
    achieved = FLOP_PER_PAIR * pairs / t_launch
    traffic = read_traffic(ROOT / "profiles" / "force_mixed_n262144_pmc.txt")
    roofline = {
        # the compute roof (of the two the schema names, hbm | mfma); the pair
        # kernel issues FP32 vector (VALU) instructions, no MFMA, and 157.3 TFLOP/s
        # is MI355X's FP32 peak
        "bound": "mfma",
        "bound_detail": "compute roof: FP32 vector ALU issue (VALU); no MFMA is used, the pair sum "
                        "is not a dense contraction; peak = MI355X FP32 157.3 TFLOP/s",
        "achieved": achieved / 1e12,
        "peak": PEAK_FP32 / 1e12,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32,
        "traffic": traffic,
        "kernel": (f"tile {m['tile']}: sym_force_kernel<K> (each unordered pair once, FP32, compensated) "
                   "+ sym_reduce_kernel<EpiMixed> (FP64 sum of partials, FP64 kick-drift, FP32 snapshot); "
                   "avg_launch_ms is the two kernels together, one step"
                   if m["tile"].startswith("sym") else
                   f"tile {m['tile']}: force_kernel<K,S,W,COMP=1,EpiMixed> (fused force, FP64 kick-drift, "
                   "FP32 snapshot); avg_launch_ms is one step"),
        "flop_per_launch": FLOP_PER_PAIR * pairs,
        "avg_launch_ms": t_launch * 1e3,
    }

    # ---- the targets of the workload, pass/fail ---------------------------------
    def target(value, op, bound, what):
        ok = None if value is None else (value >= bound if op == ">=" else value <= bound)
        return {"value": value, "target": f"{op} {bound}", "pass": ok, "what": what}

    adj_ratio = fwd_adj.get("ratio_vs_forward") if isinstance(fwd_adj, dict) else None
    targets = {
        "fwd_frac": target(achieved / PEAK_FP32, ">=", 0.6,
                           "forward step, 20 FLOP x N^2 per step / step time / 157.3 TFLOP/s"),
        "fwd_adj_ratio": target(adj_ratio, ">=", 0.4,
                                f"forward+adjoint forward-equivalent rate / forward rate, N={n}, "
                                f"{K} steps, checkpoint every 10"),
        "final_pos_relerr": target(acc["mixed"]["final_pos_relerr"], "<=", 1e-6,
                                   f"mixed path vs the FP64 path after {K} steps"),
        "grad_fd_relerr": target(fdchk.get("max_relerr") if isinstance(fdchk, dict) else None, "<=", 1e-5,
                                 "mixed adjoint gradients vs FP64 central differences, 64 coordinates, "
                                 "N=16384 cube, 50 steps (max over coordinates)"),
    }
    for k, t in targets.items():
        log(f"target {k}: {t['value']} {t['target']} -> {'pass' if t['pass'] else 'MISS'}")

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("BENCH_CPU_THREADS", min(16, os.cpu_count() or 1)))
        cpu = cpu_baseline(c, n, threads, args.cpu_rows)
        log(f"cpu baseline: {cpu['value']:.4e} interactions/s on {threads} threads")

    out = {
        "metric": METRIC,
        "value": m["interactions_per_s"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": m["wall_ms"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Plummer sphere)",
        "config": {
            "workload": f"nbody all-pairs force + kick-drift, N={n} Plummer (a=1, M=1), eps=0.002, "
                        f"dt=2.5e-4, {K} steps, FP32 compensated pairs + FP64 state",
            "n": n, "eps": 0.002, "dt": 2.5e-4, "steps": K, "precision": "mixed",
            "masses": "equal (1/N), as the reference's initial conditions (mass 1 per body)",
            "parallelism": "single GPU",
        },
        "unequal_masses": unequal,
        "vs_plain_fp32": {
            "fp32_path": "PREC_FP32: plain FP32 pair sums (uncompensated block partials, FP32 "
                         "reduction of the partials), FP32 state",
            "fp32_tile": results["fp32"]["tile"],
            "fp32_interactions_per_s": results["fp32"]["interactions_per_s"],
            "mixed_over_fp32": m["interactions_per_s"] / results["fp32"]["interactions_per_s"],
            "fp32_accel_relerr": acc["fp32"]["accel_relerr"],
            "mixed_accel_relerr": acc["mixed"]["accel_relerr"],
            "fp32_final_pos_relerr": acc["fp32"]["final_pos_relerr"],
            "mixed_final_pos_relerr": acc["mixed"]["final_pos_relerr"],
        },
        "targets": targets,
        "variants": results,
        "accuracy_vs_fp64": acc,
        "plummer65536_vs_fp64": cfg2,
        "energy": energy,
        "fwd_adjoint": fwd_adj,
        "cube_logm16384_fwd_adjoint": cfg4,
        "grad_fd_check": fdchk,
        "reference_arithmetic": ref_arith,
        "batched": batch,
        "slice_per_gpu": slc,
        "large_n": big,
        "small_systems": small,
        "fraction_of_fp32_peak": achieved / PEAK_FP32,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device": {"name": name, "cu": cus, "clock_khz": clk},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
