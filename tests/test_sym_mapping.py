"""CPU check of the symmetric kernels' work mapping (utc_amd/csrc/sym.hip,
sym_vjp.hip, sym_reduce.h), restated from sym_wave / sym_split_sum: wave block w, split r of R
takes the sub-units [Q*r/R, Q*(r+1)/R) of Q = K (NW/2 + 1), sub-unit q being
source block q % K of unit q / K (unit 0: the wave block's own sources; for
K <= 4 own block kb meets target blocks k < kb both sides and block kb target
side only (sym_own), for K = 8 every target block target side only; unit 1 + d:
the wave block w + 1 + d mod NW, both sides, the block half way round only
from the lower half when NW is even); the reduction of
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
                if u == 0 and K <= 4:
                    for k in range(kb):         # own block kb, both sides with k < kb
                        add((w, k, w, kb))
                        add((w, kb, w, k))
                    add((w, kb, w, kb))         # and target side with itself
                    continue
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


def reduce_lanes(bodies):
    """sym_reduce_lanes (utc_amd/csrc/sym_reduce.h)."""
    T = 1
    while T < 16 and bodies * T < (1 << 19):
        T *= 2
    return T


@pytest.mark.parametrize("T", [1, 2, 4, 8, 16])
def test_reduction_lanes_cover_every_body(T):
    """sym_reduce_lane<T> over the launch grid ceil(n T / 256) x 256: every body
    below n gets T lanes t = 0..T-1 of one wave, the butterfly offsets
    32, 16, .., 64/T flip exactly the bits of t, and clamped lanes read body
    n - 1 (a valid record) while writing nothing."""
    for n in (1, 5, 63, 64, 65, 1000, 4097):
        blocks = (n * T + 255) // 256
        seen = {}
        for g in range(blocks * 256):
            lane = g & 63
            t, j = divmod(lane, 64 // T)
            body = (g >> 6) * (64 // T) + j
            read = min(body, n - 1)
            assert 0 <= read < n
            if body < n:
                seen.setdefault(body, []).append((g >> 6, t))
        assert sorted(seen) == list(range(n))
        for body, lanes in seen.items():
            assert sorted(t for _, t in lanes) == list(range(T))
            assert len({wave for wave, _ in lanes}) == 1
        offs = []
        off = 32
        while off >= 64 // T:
            offs.append(off)
            off //= 2
        assert sum(offs) == 64 - 64 // T  # together they span the t bits
    assert reduce_lanes(262144) == 2 and reduce_lanes(16384) == 16 and reduce_lanes(1 << 19) == 1


def chunk_cover(NW, R, K, rs):
    """Rotation chunks (sym_wave with SPLIT): sub-unit q of wave block w is cut
    into rs chunks c of 64/rs rotation steps; chunk c starts with lane l
    holding source (l - t0) mod 64 (t0 = 64 c / rs) and, at its step t, lane l
    meets source (l - t) mod 64; a two-sided chunk stores the reaction sums that
    lane l then holds, those of source (l - t0 - steps) mod 64, to record
    ((w NW/2 + d) rs + c) 64K + kb 64 + that source; an own-block chunk shuffles
    them home from lane (l + rot) mod 64, rot = (t0 + steps) mod 64.  Counts,
    per (target body, source body), the target-side and the reaction
    evaluations, and the reaction records written and read (sym_split_sum with
    rsl = log2 rs)."""
    half = NW // 2
    Q = (half + 1) * K
    QS = Q * rs
    steps = 64 // rs
    B = 64 * K
    side, react, writes = {}, {}, {}
    for w in range(NW):
        for r in range(R):
            for qs in range(QS * r // R, QS * (r + 1) // R):
                q, c = divmod(qs, rs)
                u, kb = divmod(q, K)
                t0 = c * steps
                if u == 0:
                    for lane in range(64):
                        for t in range(t0, t0 + steps):
                            s = w * B + kb * 64 + (lane - t) % 64
                            for k in range(kb + 1):  # slots k < kb two-sided, slot kb target side
                                i = w * B + k * 64 + lane
                                side[(i, s)] = side.get((i, s), 0) + 1
                                if k < kb:
                                    react[(s, i)] = react.get((s, i), 0) + 1
                    rot = (t0 + steps) % 64
                    for h in range(64):  # lane h receives the sums of lane (h + rot): source h
                        assert ((h + rot) % 64 - t0 - steps) % 64 == h
                    continue
                d = u - 1
                if 2 * (d + 1) == NW and w >= half:
                    continue
                v = (w + 1 + d) % NW
                for lane in range(64):
                    for t in range(t0, t0 + steps):
                        s = v * B + kb * 64 + (lane - t) % 64
                        for k in range(K):
                            i = w * B + k * 64 + lane
                            side[(i, s)] = side.get((i, s), 0) + 1
                            react[(s, i)] = react.get((s, i), 0) + 1
                    el = (lane - t0 - steps) % 64
                    rec = ((w * half + d) * rs + c) * B + kb * 64 + el
                    writes[rec] = writes.get(rec, 0) + 1
    reads = {}
    for b in range(NW * B):
        v, loc = divmod(b, B)
        for dc in range(half * rs):
            d, c = divmod(dc, rs)
            w = (v - 1 - d) % NW
            if 2 * (d + 1) == NW and w >= half:
                continue
            rec = ((w * half + d) * rs + c) * B + loc
            reads[rec] = reads.get(rec, 0) + 1
    return side, react, writes, reads


@pytest.mark.parametrize("rs", [1, 2, 4, 8])
@pytest.mark.parametrize("K,R", [(2, 1), (2, 5), (4, 3), (4, 16)])
def test_rotation_chunks_evaluate_every_ordered_pair_once(K, R, rs):
    for NW in (1, 2, 3, 4):
        side, react, writes, reads = chunk_cover(NW, R, K, rs)
        n = NW * 64 * K
        both = {}
        for key, cnt in list(side.items()) + list(react.items()):
            both[key] = both.get(key, 0) + cnt
        assert len(both) == n * n, (NW, R, K, rs)
        assert all(cnt == 1 for cnt in both.values()), (NW, R, K, rs)
        assert all(cnt == 1 for cnt in writes.values()) and reads == writes, (NW, R, K, rs)
