"""utc_amd -- MI355X (gfx950) implementation of the UTC nbody hot path.

The product is the C ABI of ``include/utcnb.h`` in ``utc_amd/lib/libutcnb.so``
(HIP kernels) and the C++ UTC task shim and drivers built beside it.  This
Python module is plumbing for the tests and the benchmark: a ctypes binding of
that ABI.  It never falls back to another implementation: if the library is
missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
LIB_PATH = ROOT / "lib" / "libutcnb.so"

# ---- constants mirrored from include/utcnb.h --------------------------------
OK, EINVAL, EHIP, ENOMEM, ENODEV, ESTATE = 0, -1, -2, -3, -4, -5
PREC_REFERENCE, PREC_FP32, PREC_MIXED, PREC_FP64 = 0, 1, 2, 3
KICK_REFERENCE, KICK_PHYSICAL = 0, 1
F32, F64 = 0, 1
HOST, DEVICE = 0, 1
FLAG_NO_NEWTON = 1
FLAG_SYM_CHUNKS = 2
GEN_CUBE, GEN_CUBE_LOGM, GEN_PLUMMER, GEN_TARGETS = 0, 1, 2, 3

# Every symbol the header declares (checked by the CPU tests).
EXPORTS = [
    "utcnb_integrate_f32", "utcnb_integrate_f64", "utcnb_create", "utcnb_destroy",
    "utcnb_set_state", "utcnb_get_state", "utcnb_step", "utcnb_sync", "utcnb_accel",
    "utcnb_diagnostics", "utcnb_adjoint", "utcnb_last_kernel_ms", "utcnb_get_tile", "utcnb_batch_create",
    "utcnb_batch_destroy", "utcnb_batch_set_state", "utcnb_batch_get_state",
    "utcnb_batch_run", "utcnb_batch_get_tile", "utcnb_autotune", "utcnb_tile_count", "utcnb_tile_name",
    "utcnb_rng_seed", "utcnb_rng_next", "utcnb_splitmix64", "utcnb_generate",
    "utcnb_last_error", "utcnb_version", "utcnb_device_info", "utcnb_accel_noalloc",
    "utcnb_integrate_slice_f32", "utcnb_integrate_slice_f64", "utcnb_step_slice", "utcnb_commit_step",
    "utcnb_position_buffer", "utcnb_accel_slice", "utcnb_slice_tile", "utcnb_share_range",
    "utcnb_share_pairs", "utcnb_share_finish", "utcnb_sym_chunks",
]


class Config(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("precision", C.c_int32),
        ("kick", C.c_int32),
        ("state_dtype", C.c_int32),
        ("flags", C.c_int32),
        ("tile", C.c_int32),
        ("eps2", C.c_double),
        ("damping", C.c_double),
        ("dt", C.c_double),
    ]


class UtcnbError(RuntimeError):
    pass


_lib = None
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def _proto(lib):
    P = C.POINTER
    sig = {
        "utcnb_integrate_f32": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, _vp]),
        "utcnb_integrate_f64": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _vp]),
        "utcnb_create": (C.c_int, [P(Config), _vp, P(_vp)]),
        "utcnb_destroy": (C.c_int, [_vp]),
        "utcnb_set_state": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int]),
        "utcnb_get_state": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int]),
        "utcnb_step": (C.c_int, [_vp, C.c_int]),
        "utcnb_sync": (C.c_int, [_vp]),
        "utcnb_accel": (C.c_int, [_vp, _vp]),
        "utcnb_diagnostics": (C.c_int, [_vp, _dp, _dp, _vp]),
        "utcnb_adjoint": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _dp, _vp, _vp, _vp]),
        "utcnb_last_kernel_ms": (C.c_int, [_vp, _dp, P(C.c_int)]),
        "utcnb_get_tile": (C.c_int, [_vp, P(C.c_int)]),
        "utcnb_batch_create": (C.c_int, [C.c_int, _vp, P(Config), _vp, P(_vp)]),
        "utcnb_batch_destroy": (C.c_int, [_vp]),
        "utcnb_batch_set_state": (C.c_int, [_vp, _vp, _vp]),
        "utcnb_batch_get_state": (C.c_int, [_vp, _vp, _vp]),
        "utcnb_batch_run": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _dp]),
        "utcnb_batch_get_tile": (C.c_int, [_vp, P(C.c_int)]),
        "utcnb_autotune": (C.c_int, [C.c_int, C.c_int, C.c_char_p, P(C.c_int), _dp]),
        "utcnb_tile_count": (C.c_int, []),
        "utcnb_tile_name": (C.c_char_p, [C.c_int]),
        "utcnb_rng_seed": (None, [C.c_uint64, P(C.c_uint64)]),
        "utcnb_rng_next": (C.c_uint64, [P(C.c_uint64)]),
        "utcnb_splitmix64": (C.c_uint64, [P(C.c_uint64)]),
        "utcnb_generate": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double, _vp, _vp, _vp]),
        "utcnb_last_error": (C.c_char_p, []),
        "utcnb_version": (C.c_char_p, []),
        "utcnb_device_info": (C.c_int, [C.c_char_p, C.c_int, P(C.c_int), P(C.c_int)]),
        "utcnb_accel_noalloc": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _vp, _vp]),
        "utcnb_integrate_slice_f32": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                                C.c_float, C.c_int, _vp]),
        "utcnb_integrate_slice_f64": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                                C.c_double, C.c_int, _vp]),
        "utcnb_step_slice": (C.c_int, [_vp, C.c_int, C.c_int]),
        "utcnb_commit_step": (C.c_int, [_vp]),
        "utcnb_position_buffer": (C.c_int, [_vp, C.c_int, P(_vp), P(C.c_int)]),
        "utcnb_accel_slice": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
        "utcnb_slice_tile": (C.c_int, [_vp, C.c_int, P(C.c_int)]),
        "utcnb_share_range": (C.c_int, [_vp, C.c_int, C.c_int, P(C.c_int), P(C.c_int)]),
        "utcnb_share_pairs": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
        "utcnb_share_finish": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
        "utcnb_sym_chunks": (C.c_int, [_vp, P(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args


def lib():
    """The loaded libutcnb.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise UtcnbError(
                f"{LIB_PATH} is missing: build it with `python -m utc_amd.build` "
                "(there is no fallback implementation)")
        _lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
        _proto(_lib)
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != OK:
        msg = lib().utcnb_last_error().decode(errors="replace")
        raise UtcnbError(f"{what} failed with status {rc}: {msg}")


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def generate(kind: int, n: int, seed: int, p0=0.0, p1=0.0, p2=0.0):
    """Run a library generator; returns (pos4, vel4) or targets [n,3] (FP64)."""
    L = lib()
    if kind == GEN_TARGETS:
        out = np.zeros((n, 3), np.float64)
        check(L.utcnb_generate(kind, n, seed, p0, p1, p2, None, None, ptr(out)), "generate")
        return out
    pos = np.zeros((n, 4), np.float64)
    vel = np.zeros((n, 4), np.float64)
    check(L.utcnb_generate(kind, n, seed, p0, p1, p2, ptr(pos), ptr(vel), None), "generate")
    return pos, vel


class System:
    """One nbody system on the GPU (a handle of the C ABI)."""

    def __init__(self, n, precision, eps2, dt, damping=1.0, kick=KICK_PHYSICAL,
                 state_dtype=F64, flags=0, tile=0, stream=None):
        self.cfg = Config(n, precision, kick, state_dtype, flags, tile, eps2, damping, dt)
        self.n = n
        self.h = C.c_void_p()
        check(lib().utcnb_create(C.byref(self.cfg), stream, C.byref(self.h)), "utcnb_create")

    def close(self):
        if self.h:
            lib().utcnb_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, pos4: np.ndarray, vel4: np.ndarray | None = None):
        dt = F32 if pos4.dtype == np.float32 else F64
        pos4 = np.ascontiguousarray(pos4)
        if vel4 is not None:
            vel4 = np.ascontiguousarray(vel4, dtype=pos4.dtype)
        check(lib().utcnb_set_state(self.h, ptr(pos4), None if vel4 is None else ptr(vel4), dt, HOST),
              "utcnb_set_state")

    def get_state(self, dtype=np.float64):
        pos = np.zeros((self.n, 4), dtype)
        vel = np.zeros((self.n, 4), dtype)
        check(lib().utcnb_get_state(self.h, ptr(pos), ptr(vel), F32 if dtype == np.float32 else F64, HOST),
              "utcnb_get_state")
        return pos, vel

    def step(self, nsteps: int):
        check(lib().utcnb_step(self.h, nsteps), "utcnb_step")

    def sync(self):
        check(lib().utcnb_sync(self.h), "utcnb_sync")

    def accel(self):
        a = np.zeros((self.n, 3), np.float64)
        check(lib().utcnb_accel(self.h, ptr(a)), "utcnb_accel")
        return a

    def diagnostics(self):
        k, w = C.c_double(), C.c_double()
        p = np.zeros(3, np.float64)
        check(lib().utcnb_diagnostics(self.h, C.byref(k), C.byref(w), ptr(p)), "utcnb_diagnostics")
        return k.value, w.value, p

    def step_slice(self, start: int, count: int):
        check(lib().utcnb_step_slice(self.h, start, count), "utcnb_step_slice")

    def commit_step(self):
        check(lib().utcnb_commit_step(self.h), "utcnb_commit_step")

    def position_buffer(self, which: int = 0):
        """(device address, dtype) of the current (0) or next (1) position buffer."""
        p, dt = C.c_void_p(), C.c_int()
        check(lib().utcnb_position_buffer(self.h, which, C.byref(p), C.byref(dt)), "utcnb_position_buffer")
        return p.value, dt.value

    def accel_slice(self, start: int, count: int):
        a = np.zeros((count, 3), np.float64)
        check(lib().utcnb_accel_slice(self.h, start, count, ptr(a)), "utcnb_accel_slice")
        return a

    def share_range(self, shares: int, share: int):
        """(start, count) of the bodies of share `share` of `shares` (symmetric tiles)."""
        a, b = C.c_int(), C.c_int()
        check(lib().utcnb_share_range(self.h, shares, share, C.byref(a), C.byref(b)), "utcnb_share_range")
        return a.value, b.value

    def share_pairs(self, shares: int, share: int, react3_ptr: int):
        check(lib().utcnb_share_pairs(self.h, shares, share, C.c_void_p(react3_ptr)), "utcnb_share_pairs")

    def share_finish(self, shares: int, share: int, react_sum3_ptr: int):
        check(lib().utcnb_share_finish(self.h, shares, share, C.c_void_p(react_sum3_ptr)), "utcnb_share_finish")

    def sym_chunks(self):
        """Chunks of wave blocks the symmetric steps run in (0: one launch)."""
        g = C.c_int()
        check(lib().utcnb_sym_chunks(self.h, C.byref(g)), "utcnb_sym_chunks")
        return g.value

    def slice_tile(self, count: int):
        """(tile id, name) that the slice calls run for `count` targets."""
        t = C.c_int()
        check(lib().utcnb_slice_tile(self.h, count, C.byref(t)), "utcnb_slice_tile")
        return t.value, (tiles()[t.value - 1] if t.value > 0 else "")

    def last_kernel_ms(self):
        ms, cnt = C.c_double(), C.c_int()
        check(lib().utcnb_last_kernel_ms(self.h, C.byref(ms), C.byref(cnt)), "utcnb_last_kernel_ms")
        return ms.value, cnt.value

    def tile(self):
        """(tile id, tile name) of the force kernel this system runs; (0, "")
        for the REFERENCE and FP64 precisions."""
        t = C.c_int()
        check(lib().utcnb_get_tile(self.h, C.byref(t)), "utcnb_get_tile")
        return t.value, (tiles()[t.value - 1] if t.value > 0 else "")

    def adjoint(self, nsteps: int, ckpt_every: int, target3: np.ndarray):
        n = self.n
        loss = C.c_double()
        dx = np.zeros((n, 3), np.float64)
        dv = np.zeros((n, 3), np.float64)
        dm = np.zeros(n, np.float64)
        t = np.ascontiguousarray(target3, dtype=np.float64)
        check(lib().utcnb_adjoint(self.h, nsteps, ckpt_every, ptr(t), C.byref(loss), ptr(dx), ptr(dv), ptr(dm)),
              "utcnb_adjoint")
        return loss.value, dx, dv, dm


def partition(n: int, parts: int):
    """[(start, count)] of the reference's body partition over `parts` shares:
    the first n % parts shares take one body more
    (benchapps/nbody/utc-snode-mgpu/mgpu/body_task_mgpu.cu:41-49)."""
    per, extra = divmod(n, parts)
    out = []
    for r in range(parts):
        cnt = per + 1 if r < extra else per
        start = r * (per + 1) if r < extra else r * per + extra
        out.append((start, cnt))
    return out


def device_info():
    name = C.create_string_buffer(64)
    cu, clk = C.c_int(), C.c_int()
    check(lib().utcnb_device_info(name, 64, C.byref(cu), C.byref(clk)), "utcnb_device_info")
    return name.value.decode(), cu.value, clk.value


def autotune(n: int, precision: int = PREC_MIXED, cache_path=None):
    """(tile id, ms per step) of the fastest force-kernel tile for n bodies
    (utcnb_autotune); cache_path: optional text cache file."""
    t, ms = C.c_int(), C.c_double()
    path = None if cache_path is None else str(cache_path).encode()
    check(lib().utcnb_autotune(n, precision, path, C.byref(t), C.byref(ms)), "utcnb_autotune")
    return t.value, ms.value


def tiles():
    L = lib()
    return [L.utcnb_tile_name(t).decode() for t in range(1, L.utcnb_tile_count() + 1)]


class Batch:
    """Many independent systems in one device allocation (utcnb_batch_*)."""

    def __init__(self, sizes, eps2, dt, damping=1.0, kick=KICK_PHYSICAL, tile=0, stream=None):
        self.sizes = [int(x) for x in sizes]
        self.total = sum(self.sizes)
        self.cfg = Config(0, PREC_MIXED, kick, F64, 0, tile, eps2, damping, dt)
        arr = (C.c_int32 * len(self.sizes))(*self.sizes)
        self.h = C.c_void_p()
        check(lib().utcnb_batch_create(len(self.sizes), arr, C.byref(self.cfg), stream, C.byref(self.h)),
              "utcnb_batch_create")

    def close(self):
        if self.h:
            lib().utcnb_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, pos4, vel4):
        p = np.ascontiguousarray(pos4, np.float64)
        v = np.ascontiguousarray(vel4, np.float64)
        check(lib().utcnb_batch_set_state(self.h, ptr(p), ptr(v)), "utcnb_batch_set_state")

    def get_state(self):
        p = np.zeros((self.total, 4))
        v = np.zeros((self.total, 4))
        check(lib().utcnb_batch_get_state(self.h, ptr(p), ptr(v)), "utcnb_batch_get_state")
        return p, v

    BATCH_TILES = {0: "not tuned yet", 1: "one-pass, 4 waves per item", 2: "one-pass, 8 waves per item",
                   3: "one-pass, 16 waves per item", 4: "symmetric, persistent waves on a work counter",
                   5: "symmetric, one wave per item"}

    def tile(self):
        t = C.c_int()
        check(lib().utcnb_batch_get_tile(self.h, C.byref(t)), "utcnb_batch_get_tile")
        return t.value

    def kernel_name(self):
        return self.BATCH_TILES[self.tile()]

    def run(self, nsteps, mode=0, ckpt_every=0, target3=None):
        """Returns ms, or (ms, losses, dx, dv, dm) when target3 is given."""
        ms = C.c_double()
        if target3 is None:
            check(lib().utcnb_batch_run(self.h, nsteps, mode, 0, None, None, None, None, None, C.byref(ms)),
                  "utcnb_batch_run")
            return ms.value
        t = np.ascontiguousarray(target3, np.float64)
        losses = np.zeros(len(self.sizes))
        dx = np.zeros((self.total, 3)); dv = np.zeros((self.total, 3)); dm = np.zeros(self.total)
        check(lib().utcnb_batch_run(self.h, nsteps, mode, ckpt_every, ptr(t), ptr(losses), ptr(dx), ptr(dv),
                                    ptr(dm), C.byref(ms)), "utcnb_batch_run")
        return ms.value, losses, dx, dv, dm
