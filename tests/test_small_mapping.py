"""CPU check of the small-system kernels' mapping (utc_amd/csrc/force_small.h),
restated from small_force_kernel: workgroup b owns targets b*TW .. b*TW+TW-1
(TW = TL*A); lane l = t + TL*g of wave w holds targets t + TL*a (a < A) and sums
sources split + SPL*i (split = w*G + g, G = 64/TL, SPL = G*W, i < n_pad/SPL);
the butterfly over xor offsets TL, 2TL, .., 32 combines the G lanes of a
target; thread j < TW of the workgroup runs target j's epilogue."""
import pytest

TILES = [(1, 4, 8, 2048), (1, 8, 8, 2048), (1, 8, 16, 8192), (1, 16, 16, 4096), (1, 16, 16, 8192),
         (2, 16, 16, 8192)]


@pytest.mark.parametrize("A,TL,W,CAP", TILES)
def test_every_target_meets_every_source_once(A, TL, W, CAP):
    G, TW = 64 // TL, TL * A
    SPL = G * W
    assert 1024 % SPL == 0  # n_pad is a multiple of 1024
    for n_pad in sorted({1024, CAP}):
        hits = {}
        for w in range(W):
            for lane in range(64):
                t, g = lane % TL, lane // TL
                split = w * G + g
                for a in range(A):
                    tgt = a * TL + t
                    assert 0 <= tgt < TW
                    for i in range(n_pad // SPL):
                        src = split + i * SPL
                        assert src < n_pad <= CAP
                        hits[(tgt, src)] = hits.get((tgt, src), 0) + 1
        assert len(hits) == TW * n_pad and set(hits.values()) == {1}


@pytest.mark.parametrize("A,TL,W,CAP", TILES)
def test_butterfly_combines_the_lanes_of_a_target(A, TL, W, CAP):
    offs = []
    off = TL
    while off < 64:
        offs.append(off)
        off <<= 1
    for lane in range(64):
        # the lanes reached by xor with any subset of the offsets: exactly the
        # lanes with the same t (the same targets)
        reach = {lane}
        for o in offs:
            reach |= {x ^ o for x in reach}
        assert reach == {x for x in range(64) if x % TL == lane % TL}
