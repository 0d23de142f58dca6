"""GPU test of utcnb_accel_noalloc, the field entry point that allocates no
device memory (its work memory is a file-scope device array of the library).

Two calls in a row -- the second on fewer bodies, reusing the scratch the
first one left -- must give, bit for bit, the field utcnb_accel gives for a
handle of the same precision, tile and flags; sizes whose work memory exceeds
the scratch are refused with UTCNB_EINVAL and a message.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import utc_amd as U
from nbcases import plummer

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc_field(pos32, eps2, prec, tile, flags):
    import torch
    n = len(pos32)
    d_pos = torch.from_numpy(np.ascontiguousarray(pos32)).cuda()
    d_acc = torch.full((n, 3), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = U.lib().utcnb_accel_noalloc(C.c_void_p(d_pos.data_ptr()), n, eps2, prec, tile, flags,
                                     C.c_void_p(d_acc.data_ptr()), None)
    U.check(rc, "utcnb_accel_noalloc")
    torch.cuda.synchronize()
    return d_acc.cpu().numpy()


def handle_field(pos32, eps2, prec, tile, flags):
    n = len(pos32)
    with U.System(n, prec, eps2, 1e-3, flags=flags, tile=tile) as s:
        s.set_state(pos32)
        return s.accel(), s.tile()


@pytest.mark.parametrize("prec", [U.PREC_MIXED, U.PREC_FP32])
@pytest.mark.parametrize("tile_name", ["default", "sym4r64", "sym2r32", "k2s8"])
def test_noalloc_field_equals_handle_field(prec, tile_name):
    eps2 = 0.005 ** 2
    flags = U.FLAG_NO_NEWTON if tile_name != "sym2r32" else 0
    for n in (12000, 3001):  # the second call: fewer bodies, the scratch holds the first call's data
        pos32 = plummer(n, seed=n)["pos"].astype(np.float32)
        tile = 0 if tile_name == "default" else U.tiles().index(tile_name) + 1
        ref, (tid, tname) = handle_field(pos32, eps2, prec, tile, flags)
        say(f"case: noalloc prec={prec} tile={tname} n={n}")
        got = noalloc_field(pos32, eps2, prec, tid, flags)
        assert np.array_equal(got, ref), (n, tname, float(np.abs(got - ref).max()))


def test_noalloc_single_body():
    """A lone body: its field is zero, as utcnb_accel gives."""
    pos32 = np.array([[0.3, -0.2, 0.1, 1.0]], np.float32)
    say("case: noalloc n=1")
    ref, (tid, _) = handle_field(pos32, 1e-4, U.PREC_MIXED, 0, 0)
    got = noalloc_field(pos32, 1e-4, U.PREC_MIXED, tid, 0)
    assert np.array_equal(got, ref) and not got.any()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    import torch
    say("case: noalloc size limit")
    L = U.lib()
    d = torch.zeros(4, dtype=torch.float32, device="cuda")
    a = torch.zeros(3, dtype=torch.float64, device="cuda")
    # a symmetric tile at 65536 bodies needs well over 64 MiB of partial sums
    tile = U.tiles().index("sym8r64") + 1
    rc = L.utcnb_accel_noalloc(C.c_void_p(d.data_ptr()), 65536, 1e-4, U.PREC_MIXED, tile, 0,
                               C.c_void_p(a.data_ptr()), None)
    assert rc == U.EINVAL
    assert b"64 MiB" in L.utcnb_last_error()
    rc = L.utcnb_accel_noalloc(C.c_void_p(d.data_ptr()), 1 << 23, 1e-4, U.PREC_MIXED, 0, 0,
                               C.c_void_p(a.data_ptr()), None)
    assert rc == U.EINVAL
    assert L.utcnb_accel_noalloc(C.c_void_p(d.data_ptr()), 16, 1e-4, U.PREC_FP64, 0, 0,
                                 C.c_void_p(a.data_ptr()), None) == U.EINVAL
