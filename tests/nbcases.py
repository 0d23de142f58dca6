"""Shared input builders for the tests and the benchmark (seeded workloads):

  cube4096        N=4096 uniform cube [-1,1]^3, v ~ N(0, 0.01^2), m = 1/N, eps = 0.01,
                  dt = 1e-3, 10 steps, damping 1, reference kick (as the repo runs it)
  plummer65536    N=65536 Plummer, eps = 0.005, dt = 5e-4, 20 steps, physical kick
  plummer262144   N=262144 Plummer, eps = 0.002, dt = 2.5e-4, 10 steps, physical kick
  cube_logm16384  N=16384 uniform cube, m ~ exp(U[ln .5, ln 2])/N, eps = 0.01,
                  dt = 1e-3, 50 steps, targets ~ N(0, 0.5^2), checkpoint every 10
  batch256_sizes  256 Plummer systems, n ~ round(exp(U[ln 256, ln 8192])), eps = 0.01,
                  dt = 1e-3, 20 steps
"""
from __future__ import annotations

import numpy as np

import utc_amd as U

SEED = 0
TARGET_SEED_XOR = 0x5EED7A56E7


def cube4096(n=4096, seed=SEED):
    pos, vel = U.generate(U.GEN_CUBE, n, seed, 0.01)
    return dict(pos=pos, vel=vel, eps=0.01, dt=1e-3, steps=10, damping=1.0,
                kick=U.KICK_REFERENCE)


def plummer(n, seed=SEED, eps=0.005, dt=5e-4, steps=20):
    pos, vel = U.generate(U.GEN_PLUMMER, n, seed)
    return dict(pos=pos, vel=vel, eps=eps, dt=dt, steps=steps, damping=1.0,
                kick=U.KICK_PHYSICAL)


def plummer65536(n=65536, seed=SEED):
    return plummer(n, seed, 0.005, 5e-4, 20)


def plummer262144(n=262144, seed=SEED):
    return plummer(n, seed, 0.002, 2.5e-4, 10)


def cube_logm16384(n=16384, seed=SEED, steps=50):
    pos, vel = U.generate(U.GEN_CUBE_LOGM, n, seed, 0.01, 0.5, 2.0)
    tgt = U.generate(U.GEN_TARGETS, n, seed ^ TARGET_SEED_XOR, 0.5)
    return dict(pos=pos, vel=vel, target=tgt, eps=0.01, dt=1e-3, steps=steps, damping=1.0,
                kick=U.KICK_PHYSICAL, ckpt=10)


def batch256_sizes(nsys=256, seed=SEED):
    st = np.zeros(4, np.uint64)
    import ctypes as C
    s = (C.c_uint64 * 4)()
    U.lib().utcnb_rng_seed(seed ^ 0xBA7C4, s)
    out = []
    for _ in range(nsys):
        u = (U.lib().utcnb_rng_next(s) >> 11) * 2.0 ** -53
        out.append(int(round(np.exp(np.log(256) + (np.log(8192) - np.log(256)) * u))))
    return out


def relnorm(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def mass_pattern(pos, pattern, seed=5):
    """Masses (total 1) for the symmetric kernels' two pair forms: blocks whose
    targets and sources share one mass drop the mass factors."""
    n = len(pos)
    p = pos.copy()
    rng = np.random.default_rng(seed)
    if pattern == "equal":  # one mass, not a power of two: every block uniform
        m = np.full(n, 1.0 / n)
    elif pattern == "one_heavy":  # uniform except the blocks that hold body n/3
        m = np.full(n, 1.0)
        m[n // 3] = 50.0
    elif pattern == "few_heavy":  # a scatter of different masses: mixed blocks
        m = np.full(n, 1.0)
        m[rng.choice(n, 12, replace=False)] = rng.uniform(2, 20, 12)
    else:  # log-uniform: no uniform block
        m = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n))
    p[:, 3] = m / m.sum()
    return p
