"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
sequential nbody compiled under _ref/ (when present).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke entry point and the
CPU-baseline leg of the benchmark, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_seq.so"
REF_BIN = HERE / "_ref" / "nbody_main_seq"

_o = None
_r = None
_vp = C.c_void_p
_f = C.c_float
_d = C.c_double


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def lib():
    global _o
    if _o is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        o = C.CDLL(str(LIB))
        o.orc_gravitation_f32.argtypes = [C.c_int, _vp, _f, _vp]
        o.orc_gravitation_f64.argtypes = [C.c_int, _vp, _d, _vp]
        o.orc_gravitation_rows_f32.argtypes = [C.c_int, _vp, _f, C.c_int, C.c_int, _vp, C.c_int]
        o.orc_integrate_f32.argtypes = [C.c_int, _vp, _vp, _vp, _f, _f]
        o.orc_integrate_f64.argtypes = [C.c_int, _vp, _vp, _vp, _d, _d]
        o.orc_run_f32.argtypes = [C.c_int, C.c_int, _vp, _vp, _f, _f, _f, C.c_int]
        o.orc_run_f64.argtypes = [C.c_int, C.c_int, _vp, _vp, _d, _d, _d]
        o.orc_run_f64_mt.argtypes = [C.c_int, C.c_int, _vp, _vp, _d, _d, _d, C.c_int]
        o.orc_srand.argtypes = [C.c_uint]
        o.orc_randomize_f32.argtypes = [C.c_int, _vp, _vp, _vp, _f, _f, C.c_int, C.c_int]
        o.orc_randomize_f64.argtypes = [C.c_int, _vp, _vp, _vp, _f, _f, C.c_int, C.c_int]
        o.orc_demo_params.argtypes = [C.c_int, C.c_int, _vp]
        o.orc_write_frames_f32.argtypes = [C.c_char_p, _vp, C.c_int, C.c_int, _f]
        o.orc_write_frames_f64.argtypes = [C.c_char_p, _vp, C.c_int, C.c_int, _f]
        o.orc_accel_f64.argtypes = [C.c_int, _vp, _d, _vp]
        o.orc_accel_ld.argtypes = [C.c_int, _vp, _d, _vp]
        o.orc_accel_rows_f64.argtypes = [C.c_int, _vp, _d, C.c_int, _vp, _vp, C.c_int]
        o.orc_step_f64.argtypes = [C.c_int, _vp, _vp, _d, _d, _d, C.c_int]
        o.orc_energy_f64.argtypes = [C.c_int, _vp, _vp, _d, C.POINTER(_d), C.POINTER(_d), _vp]
        o.orc_loss_f64.argtypes = [C.c_int, C.c_int, _vp, _vp, _vp, _d, _d, _d, C.c_int]
        o.orc_loss_f64.restype = _d
        o.orc_adjoint_f64.argtypes = [C.c_int, C.c_int, _vp, _vp, _vp, _d, _d, _d, C.c_int, _vp, _vp, _vp]
        o.orc_adjoint_f64.restype = _d
        _o = o
    return _o


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib():
    global _r
    if _r is None:
        r = C.CDLL(str(REF_LIB))
        r.ref_run_f32.argtypes = [C.c_int, C.c_int, _vp, _vp, _f, _f, _f]
        r.ref_run_f64.argtypes = [C.c_int, C.c_int, _vp, _vp, _d, _d, _d]
        r.ref_randomize_f32.argtypes = [C.c_int, _vp, _vp, _vp, _f, _f, C.c_int, C.c_int]
        r.ref_randomize_f64.argtypes = [C.c_int, _vp, _vp, _vp, _f, _f, C.c_int, C.c_int]
        r.ref_srand.argtypes = [C.c_uint]
        _r = r
    return _r


# ---- convenience wrappers ------------------------------------------------------
def demo_params(preset: int, n: int):
    out = np.zeros(5, np.float32)
    lib().orc_demo_params(preset, n, _ptr(out))
    return dict(timestep=out[0], clusterScale=out[1], velocityScale=out[2], softening=out[3],
                damping=out[4])


def randomize(config: int, n: int, clusterScale, velocityScale, dtype=np.float32, seed=1,
              color=True):
    """The reference's SHELL/RANDOM/EXPAND initial conditions (glibc rand())."""
    pos = np.zeros((n, 4), dtype)
    vel = np.zeros((n, 4), dtype)
    col = np.zeros((n, 4), np.float32) if color else None
    lib().orc_srand(seed)
    f = lib().orc_randomize_f32 if dtype == np.float32 else lib().orc_randomize_f64
    f(config, _ptr(pos), _ptr(vel), None if col is None else _ptr(col), _f(clusterScale),
      _f(velocityScale), n, 1)
    return pos, vel


def run_f32(pos4, vel4, steps, eps2, damping, dt, nthreads=1):
    p = np.ascontiguousarray(pos4, np.float32).copy()
    v = np.ascontiguousarray(vel4, np.float32).copy()
    lib().orc_run_f32(len(p), steps, _ptr(p), _ptr(v), _f(eps2), _f(damping), _f(dt), nthreads)
    return p, v


def run_f64(pos4, vel4, steps, eps2, damping, dt, nthreads=1):
    p = np.ascontiguousarray(pos4, np.float64).copy()
    v = np.ascontiguousarray(vel4, np.float64).copy()
    lib().orc_run_f64_mt(len(p), steps, _ptr(p), _ptr(v), _d(eps2), _d(damping), _d(dt), nthreads)
    return p, v


def gravitation_f32(pos4, eps2):
    p = np.ascontiguousarray(pos4, np.float32)
    f = np.zeros((len(p), 3), np.float32)
    lib().orc_gravitation_f32(len(p), _ptr(p), _f(eps2), _ptr(f))
    return f


def gravitation_rows_f32(pos4, eps2, i0, i1, nthreads=1):
    p = np.ascontiguousarray(pos4, np.float32)
    f = np.zeros((len(p), 3), np.float32)
    lib().orc_gravitation_rows_f32(len(p), _ptr(p), _f(eps2), i0, i1, _ptr(f), nthreads)
    return f


def gravitation_f64(pos4, eps2):
    p = np.ascontiguousarray(pos4, np.float64)
    f = np.zeros((len(p), 3), np.float64)
    lib().orc_gravitation_f64(len(p), _ptr(p), _d(eps2), _ptr(f))
    return f


def accel_f64(pos4, eps2):
    p = np.ascontiguousarray(pos4, np.float64)
    a = np.zeros((len(p), 3), np.float64)
    lib().orc_accel_f64(len(p), _ptr(p), _d(eps2), _ptr(a))
    return a


def accel_rows_f64(pos4, eps2, rows, nthreads=8):
    """Rows `rows` of accel_f64 (bit for bit), threaded."""
    p = np.ascontiguousarray(pos4, np.float64)
    r = np.ascontiguousarray(rows, np.int32)
    a = np.zeros((len(r), 3), np.float64)
    if lib().orc_accel_rows_f64(len(p), _ptr(p), _d(eps2), len(r), _ptr(r), _ptr(a), nthreads) != 0:
        raise ValueError("row index out of range")
    return a


def accel_ld(pos4, eps2):
    p = np.ascontiguousarray(pos4, np.float64)
    a = np.zeros((len(p), 3), np.float64)
    lib().orc_accel_ld(len(p), _ptr(p), _d(eps2), _ptr(a))
    return a


def steps_f64(pos4, vel4, steps, eps2, dt, damping=1.0, kick=1):
    p = np.ascontiguousarray(pos4, np.float64).copy()
    v = np.ascontiguousarray(vel4, np.float64).copy()
    for _ in range(steps):
        lib().orc_step_f64(len(p), _ptr(p), _ptr(v), _d(eps2), _d(dt), _d(damping), kick)
    return p, v


def energy_f64(pos4, vel4, eps2):
    p = np.ascontiguousarray(pos4, np.float64)
    v = np.ascontiguousarray(vel4, np.float64)
    k, w = _d(), _d()
    mom = np.zeros(3, np.float64)
    lib().orc_energy_f64(len(p), _ptr(p), _ptr(v), _d(eps2), C.byref(k), C.byref(w), _ptr(mom))
    return k.value, w.value, mom


def loss_f64(pos4, vel4, target3, steps, eps2, dt, damping=1.0, kick=1):
    p = np.ascontiguousarray(pos4, np.float64)
    v = np.ascontiguousarray(vel4, np.float64)
    t = np.ascontiguousarray(target3, np.float64)
    return lib().orc_loss_f64(len(p), steps, _ptr(p), _ptr(v), _ptr(t), eps2, dt, damping, kick)


def adjoint_f64(pos4, vel4, target3, steps, eps2, dt, damping=1.0, kick=1):
    n = len(pos4)
    p = np.ascontiguousarray(pos4, np.float64)
    v = np.ascontiguousarray(vel4, np.float64)
    t = np.ascontiguousarray(target3, np.float64)
    dx = np.zeros((n, 3)); dv = np.zeros((n, 3)); dm = np.zeros(n)
    L = lib().orc_adjoint_f64(n, steps, _ptr(p), _ptr(v), _ptr(t), eps2, dt, damping, kick,
                              _ptr(dx), _ptr(dv), _ptr(dm))
    return L, dx, dv, dm


def write_frames(path, frames, timestep):
    frames = np.ascontiguousarray(frames)
    nf, n = frames.shape[0], frames.shape[1]
    f = lib().orc_write_frames_f32 if frames.dtype == np.float32 else lib().orc_write_frames_f64
    rc = f(str(path).encode(), _ptr(frames), nf, n, _f(timestep))
    if rc != 0:
        raise OSError(f"cannot write {path}")
