"""CPU check of the symmetric kernels' work mapping (utc_amd/csrc/sym.hip,
sym_vjp.hip), restated from sym_wave / sym_sum: wave block w, split r of R
takes work units [U*r/R, U*(r+1)/R) of U = NW/2 + 1 (unit 0: its own source
blocks, target side only; unit 1 + d: the block w + 1 + d mod NW, both sides,
the block half way round only from the lower half when NW is even), and the
reduction of body block v reads the reaction partial (w = v - 1 - d mod NW, d).
Every ordered pair of wave blocks must be evaluated exactly once, every
reaction slot written once and read once."""
import pytest


def cover(NW, R):
    half = NW // 2
    units = half + 1
    pairs, writes = {}, {}
    for w in range(NW):
        for r in range(R):
            u0, u1 = units * r // R, units * (r + 1) // R
            if u0 == 0 and u1 > 0:
                pairs[(w, w)] = pairs.get((w, w), 0) + 1
            for d in range((u0 - 1) if u0 > 0 else 0, u1 - 1):
                if 2 * (d + 1) == NW and w >= half:
                    continue
                v = (w + 1 + d) % NW
                pairs[(w, v)] = pairs.get((w, v), 0) + 1   # targets of w, sources of v
                pairs[(v, w)] = pairs.get((v, w), 0) + 1   # the reaction on v
                writes[(w, d)] = writes.get((w, d), 0) + 1
    reads = {}
    for v in range(NW):
        for d in range(half):
            w = (v - 1 - d) % NW
            if 2 * (d + 1) == NW and w >= half:
                continue
            reads[(w, d)] = reads.get((w, d), 0) + 1
    return pairs, writes, reads


@pytest.mark.parametrize("R", [1, 2, 3, 4, 9, 16, 32, 64])
def test_every_ordered_pair_of_wave_blocks_once(R):
    for NW in range(1, 70):
        pairs, writes, reads = cover(NW, R)
        assert all(pairs.get((a, b), 0) == 1 for a in range(NW) for b in range(NW)), (NW, R)
        assert len(pairs) == NW * NW
        assert all(c == 1 for c in writes.values()), (NW, R)
        assert reads == writes, (NW, R)
