"""CPU tests of the C ABI library: it loads, exports every declared symbol,
validates arguments before touching a device, and its host-side generators
follow their published definitions.  No GPU compute is called here."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import utc_amd as U

HEADER = Path(__file__).resolve().parent.parent / "include" / "utcnb.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(utcnb_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = U.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(U.EXPORTS) == names


def test_version_and_tiles():
    L = U.lib()
    assert L.utcnb_version().startswith(b"utcnb")
    assert L.utcnb_tile_count() == 37
    assert U.tiles()[8] == "k2s8" and U.tiles()[12] == "sym8r4"
    assert L.utcnb_tile_name(0) == b"" and L.utcnb_tile_name(99) == b""


def test_splitmix64_published_vector():
    # splitmix64 (Steele, Lea & Flood 2014) from state 0
    x = C.c_uint64(0)
    got = [U.lib().utcnb_splitmix64(C.byref(x)) for _ in range(4)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


def test_xoshiro256starstar_published_vector():
    # xoshiro256** (Blackman & Vigna 2018) from state {1, 2, 3, 4}
    s = (C.c_uint64 * 4)(1, 2, 3, 4)
    got = [U.lib().utcnb_rng_next(s) for _ in range(10)]
    assert got == [11520, 0, 1509978240, 1215971899390074240, 1216172134540287360,
                   607988272756665600, 16172922978634559625, 8476171486693032832,
                   10595114339597558777, 2904607092377533576]


def test_cube_generator_moments():
    n = 200000
    pos, vel = U.generate(U.GEN_CUBE, n, 0, 0.01)
    assert pos[:, :3].min() >= -1 and pos[:, :3].max() < 1
    assert abs(pos[:, :3].mean()) < 5e-3
    assert pos[:, :3].var() == pytest.approx(1 / 3, rel=1e-2)
    assert np.all(pos[:, 3] == 1.0 / n)
    assert abs(vel[:, :3].mean()) < 1e-4
    assert vel[:, :3].std() == pytest.approx(0.01, rel=1e-2)
    z = vel[:, :3].ravel() / 0.01
    assert np.mean(z ** 4) == pytest.approx(3.0, rel=3e-2)  # normal kurtosis
    assert np.all(vel[:, 3] == 1.0)
    # same seed, same stream; different seed, different stream
    p2, _ = U.generate(U.GEN_CUBE, n, 0, 0.01)
    assert np.array_equal(p2, pos)
    p3, _ = U.generate(U.GEN_CUBE, n, 1, 0.01)
    assert not np.array_equal(p3, pos)


def test_logm_generator_range_and_distribution():
    n = 100000
    pos, _ = U.generate(U.GEN_CUBE_LOGM, n, 3, 0.01, 0.5, 2.0)
    m = pos[:, 3] * n
    assert m.min() >= 0.5 and m.max() <= 2.0
    # log-uniform: ln m uniform on [ln .5, ln 2]
    u = (np.log(m) - math.log(0.5)) / (math.log(2.0) - math.log(0.5))
    assert np.mean(u) == pytest.approx(0.5, abs=5e-3)
    assert np.var(u) == pytest.approx(1 / 12, rel=2e-2)


def test_targets_generator():
    t = U.generate(U.GEN_TARGETS, 100000, 9, 0.5)
    assert t.shape == (100000, 3)
    assert t.std() == pytest.approx(0.5, rel=1e-2)


def test_plummer_mass_profile_and_virial():
    n = 20000
    pos, vel = U.generate(U.GEN_PLUMMER, n, 0)
    assert pos[:, 3].sum() == pytest.approx(1.0, rel=1e-12)
    # centre of mass and mean velocity removed
    assert np.abs(pos[:, :3].mean(0)).max() < 1e-12
    assert np.abs(vel[:, :3].mean(0)).max() < 1e-12
    r = np.sort(np.linalg.norm(pos[:, :3], axis=1))
    # cumulative mass M(<r) = r^3 / (1 + r^2)^{3/2}, truncated at mass fraction 0.999
    model = r ** 3 / (1 + r * r) ** 1.5 / 0.999
    emp = np.arange(1, n + 1) / n
    ks = np.abs(emp - model).max()
    assert ks < 1.63 / math.sqrt(n)  # 1% Kolmogorov-Smirnov critical value
    # speeds never exceed the local escape speed sqrt(2) (1 + r^2)^{-1/4}
    from oracle import oracle as O
    k, w, _ = O.energy_f64(pos, vel, 0.0)
    assert 2 * k / abs(w) == pytest.approx(1.0, abs=0.05)
    assert w == pytest.approx(-3 * math.pi / 32, rel=0.05)


def test_create_validates_before_touching_a_device():
    L = U.lib()
    h = C.c_void_p()
    bad = [
        U.Config(0, U.PREC_MIXED, 1, 1, 0, 0, 1e-4, 1.0, 1e-3),      # n < 1
        U.Config((1 << 30) + 1, U.PREC_MIXED, 1, 1, 0, 0, 1e-4, 1.0, 1e-3),  # n above 2^30
        U.Config(10, 7, 1, 1, 0, 0, 1e-4, 1.0, 1e-3),                # precision
        U.Config(10, U.PREC_MIXED, 5, 1, 0, 0, 1e-4, 1.0, 1e-3),     # kick
        U.Config(10, U.PREC_MIXED, 1, 1, 0, 0, 0.0, 1.0, 1e-3),      # eps2 = 0
        U.Config(10, U.PREC_FP32, 1, 1, 0, L.utcnb_tile_count() + 1, 1e-4, 1.0, 1e-3),  # tile
        U.Config(10, U.PREC_REFERENCE, 1, 0, 0, 0, 1e-4, 1.0, 1e-3), # reference needs its kick
        U.Config(10, U.PREC_FP64, 1, 1, 0, 0, 1e-4, 1.0, float("nan")),
    ]
    for cfg in bad:
        assert L.utcnb_create(C.byref(cfg), None, C.byref(h)) == U.EINVAL
        assert L.utcnb_last_error()
    assert L.utcnb_integrate_f32(None, None, None, 10, 0.1, 0.1, 1.0, 256, None) == U.EINVAL
    assert L.utcnb_generate(U.GEN_CUBE, 0, 0, 0, 0, 0, None, None, None) == U.EINVAL
    assert b"utcnb_generate" in L.utcnb_last_error()
    assert L.utcnb_generate(99, 10, 0, 0, 0, 0, None, None, None) == U.EINVAL


def test_create_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = U.lib()
    h = C.c_void_p()
    cfg = U.Config(10, U.PREC_MIXED, 1, 1, 0, 0, 1e-4, 1.0, 1e-3)
    assert L.utcnb_create(C.byref(cfg), None, C.byref(h)) == U.ENODEV
    assert b"device" in L.utcnb_last_error()


def test_autotune_validates_before_touching_a_device():
    L = U.lib()
    t, ms = C.c_int(), C.c_double()
    assert L.utcnb_autotune(0, U.PREC_MIXED, None, C.byref(t), C.byref(ms)) == U.EINVAL
    assert L.utcnb_autotune(100, U.PREC_FP64, None, C.byref(t), C.byref(ms)) == U.EINVAL
    assert L.utcnb_autotune(100, U.PREC_MIXED, None, None, None) == U.EINVAL
