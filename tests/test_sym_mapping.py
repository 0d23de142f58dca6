"""CPU check of the symmetric kernels' work mapping (utc_amd/csrc/sym.hip,
sym_vjp.hip), restated from sym_wave / sym_sum: wave block w, split r of R
takes the sub-units [Q*r/R, Q*(r+1)/R) of Q = K (NW/2 + 1), sub-unit q being
source block q % K of unit q / K (unit 0: the wave block's own sources, target
side only; unit 1 + d: the wave block w + 1 + d mod NW, both sides, the block
half way round only from the lower half when NW is even); the reduction of
body block v reads the reaction partials (w = v - 1 - d mod NW, d, kb).
Every ordered pair of 64-body blocks must be evaluated exactly once, every
reaction slot written once and read once."""
import pytest


def cover(NW, R, K):
    """Count, for every (target wave block a, its 64-body block ka, source wave
    block b, its block kb), how often the mapping evaluates that block pair."""
    half = NW // 2
    Q = (half + 1) * K
    hits, writes = {}, {}

    def add(key):
        hits[key] = hits.get(key, 0) + 1

    for w in range(NW):
        for r in range(R):
            for q in range(Q * r // R, Q * (r + 1) // R):
                u, kb = divmod(q, K)
                v = w
                if u > 0:
                    d = u - 1
                    if 2 * (d + 1) == NW and w >= half:
                        continue
                    v = (w + 1 + d) % NW
                    writes[(w, d, kb)] = writes.get((w, d, kb), 0) + 1
                for k in range(K):
                    add((w, k, v, kb))          # target side: w's block k from v's block kb
                    if u > 0:
                        add((v, kb, w, k))      # reaction: v's block kb from w's block k
    reads = {}
    for v in range(NW):
        for d in range(half):
            w = (v - 1 - d) % NW
            if 2 * (d + 1) == NW and w >= half:
                continue
            for kb in range(K):
                reads[(w, d, kb)] = reads.get((w, d, kb), 0) + 1
    return hits, writes, reads


@pytest.mark.parametrize("K", [1, 2, 4, 8])
@pytest.mark.parametrize("R", [1, 2, 3, 9, 32, 128])
def test_every_ordered_pair_of_blocks_once(R, K):
    for NW in list(range(1, 20)) + [32, 33, 64]:
        hits, writes, reads = cover(NW, R, K)
        assert len(hits) == (NW * K) ** 2, (NW, R, K)
        assert all(c == 1 for c in hits.values()), (NW, R, K)
        assert all(c == 1 for c in writes.values()), (NW, R, K)
        assert reads == writes, (NW, R, K)
