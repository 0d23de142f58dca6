// sym.hip -- all-pairs force evaluating every unordered pair once.
//
// The field on body i from body j and the reaction on j from i share r, d^2
// and y^3 = (|r|^2 + eps2)^{-3/2}: a_i += m_j y^3 r, a_j -= m_i y^3 r.  The
// one-pass kernels (force_tmpl.h) spend 12 VALU issue slots per ordered pair;
// here an unordered pair costs 16 (3 sub, 3 fma, 2 mul for y^3, 2 mul for the
// two mass factors, 6 fma for the two accumulations) plus, per 64-source
// rotation step, 7 wave-wide DPP moves shared by the K targets of a lane:
// (16 + 7/K)/2 slots per ordered pair, 8.4 at K = 8.  When the wave's targets
// and the block's sources all have one mass m (the reference's initial
// conditions give every body mass 1), both sides use y^3 r directly and the
// block partials are scaled by m once: 14 slots and 6 DPP moves, (14 + 6/K)/2
// = 7.4 per ordered pair.  For m a power of two the two forms are bitwise
// equal; otherwise they differ by the rounding of one product per partial.
//
// Mapping.  Bodies form wave blocks of 64*K (lane l holds K targets).  Wave
// block w takes (a) its own K blocks of 64 sources: block kb both sides
// against target slots k < kb and target side only against slot kb (K <= 4;
// at K = 8, target side only against every slot), and (b) the source blocks
// of wave blocks v = w+1 .. w+NW/2 (mod NW), both sides; for even NW the
// block half way round is taken by the lower half only, so every unordered
// pair of wave blocks is visited exactly once.  A block of 64
// sources sits one per lane and rotates around the wave with DPP wave_ror:1
// for 64 steps, carrying its reaction accumulator, so after 64 steps every
// target has met every source and each reaction sum is back in its home lane.
// R waves split a wave block's work into sub-units (u, kb): source block kb of
// unit u, where unit 0 is the wave block's own sources and unit 1 + d the wave
// block d + 1 places round; wave r takes sub-units [Q*r/R, Q*(r+1)/R) of
// Q = K (NW/2 + 1).  The waves' target-side partials and the reaction
// partials go to HBM, and a second kernel adds them per body in a fixed order
// (deterministic) and applies the integrator epilogue.
// Small systems have too few sub-units to give every SIMD several waves, and
// one wave alone on a SIMD waits on its own dependent instructions.  Tiles
// with a rotation split RS (SPLIT kernels) cut each sub-unit's 64 rotation
// steps into RS chunks of 64/RS steps, and the R waves of a wave block share
// its Q * RS chunks: chunk c starts with the source block already rotated by
// 64c/RS lanes (lane l loads source (l - 64c/RS) mod 64) and leaves the
// reaction sums rotated by 64(c+1)/RS, so each chunk stores its reactions to
// the source lane they belong to, in a reaction record of its own.
#include "force_tmpl.h"
#include "sym_items.h"
#include "sym_reduce.h"

namespace utcnb {

// Lane l receives lane l-1's value (lane 0 lane 63's).  Every lane has a
// source under wave_ror:1, so the "old" operand is never used.  Passing v as
// "old" lets the compiler rotate in place, without a v_mov that first zeroes
// a new register; which form is faster depends on K (same-device A/B runs on
// MI355X, equal masses): in place is 7.7% faster at K = 2 (N = 4096, sym2r32)
// but 3% slower at K = 8 (N = 262144, sym8r64) and 3.4% slower for the
// batched K = 4 kernel, and the same within noise for the K = 4 tiles.
template <int K>
__device__ __forceinline__ float wave_ror1(float v)
{
    const int x = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(K <= 2 ? x : 0, x, 0x13C /* wave_ror:1 */, 0xF, 0xF, false));
}

// Adds a block partial p into the running target sum: Kahan-compensated
// (COMP) or plain FP32.
template <bool COMP>
__device__ __forceinline__ void block_add(float& a, float& c, float p)
{
    if constexpr (COMP)
        kahan_add(a, c, p);
    else
        a += p;
}

// One rotation step of sym_block (statement macro: the same code in the
// fixed 64-step loop and in the chunked loop).
#define SYM_BLOCK_STEP()                                                                               \
    _Pragma("unroll")                                                                                  \
    for (int k = 0; k < K; ++k) {                                                                      \
        const float rx = sx - xi[k], ry = sy - yi[k], rz = sz - zi[k];                                 \
        const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2))); \
        const float y3 = pair_scale<NEWTON>(d2, 1.0f);                                                 \
        const float si = UNI ? y3 : sm * y3;                                                           \
        const float cj = UNI ? y3 : mi[k] * y3;                                                        \
        px[k] = __builtin_fmaf(rx, si, px[k]);                                                         \
        py[k] = __builtin_fmaf(ry, si, py[k]);                                                         \
        pz[k] = __builtin_fmaf(rz, si, pz[k]);                                                         \
        gx = __builtin_fmaf(-rx, cj, gx);                                                              \
        gy = __builtin_fmaf(-ry, cj, gy);                                                              \
        gz = __builtin_fmaf(-rz, cj, gz);                                                              \
    }                                                                                                  \
    sx = wave_ror1<K>(sx); sy = wave_ror1<K>(sy); sz = wave_ror1<K>(sz);                               \
    if constexpr (!UNI)                                                                                \
        sm = wave_ror1<K>(sm);                                                                         \
    gx = wave_ror1<K>(gx); gy = wave_ror1<K>(gy); gz = wave_ror1<K>(gz);

// A block of 64 sources (one per lane, s) against the K targets of each lane,
// both sides: the target-side block partials are added into (a, c) by
// block_add<COMP>, the reaction sums of the lane's source come back in g after
// 64 rotations.
// UNI: every target mass mi[k] and source mass equals m, so the pair terms
// drop the mass factors and the block sums are scaled by m at the end.
// SPLIT: `steps` rotation steps (a chunk of the rotation), else all 64.
template <int K, bool NEWTON, bool COMP, bool UNI, bool SPLIT = false>
__device__ __forceinline__ void sym_block(float4 s, const float (&xi)[K], const float (&yi)[K],
                                          const float (&zi)[K], const float (&mi)[K], float m, float eps2,
                                          float (&ax)[K], float (&ay)[K], float (&az)[K], float (&cx)[K],
                                          float (&cy)[K], float (&cz)[K], float& gx, float& gy, float& gz,
                                          int steps = 64)
{
    float sx = s.x, sy = s.y, sz = s.z, sm = s.w;
    float px[K], py[K], pz[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
        px[k] = py[k] = pz[k] = 0.f;
    gx = gy = gz = 0.f;
    if constexpr (SPLIT) {
        // a chunk's `steps` (a multiple of 4) in unrolled groups of 4: a loop
        // of DPP steps with a run-time trip count is not unrolled otherwise
        for (int t = 0; t < steps; t += 4)
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) {
                SYM_BLOCK_STEP();
            }
    } else {
#pragma unroll 4
        for (int t = 0; t < 64; ++t) {
            SYM_BLOCK_STEP();
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if constexpr (UNI) {
            px[k] *= m; py[k] *= m; pz[k] *= m;
        }
        block_add<COMP>(ax[k], cx[k], px[k]);
        block_add<COMP>(ay[k], cy[k], py[k]);
        block_add<COMP>(az[k], cz[k], pz[k]);
    }
    if constexpr (UNI) {
        gx *= m; gy *= m; gz *= m;
    }
}

// One rotation step of sym_own.
#define SYM_OWN_STEP()                                                                                 \
    _Pragma("unroll")                                                                                  \
    for (int k = 0; k < KB; ++k) {                                                                     \
        const float rx = sx - xi[k], ry = sy - yi[k], rz = sz - zi[k];                                 \
        const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2))); \
        const float y3 = pair_scale<NEWTON>(d2, 1.0f);                                                 \
        const float si = UNI ? y3 : sm * y3;                                                           \
        const float cj = UNI ? y3 : mi[k] * y3;                                                        \
        px[k] = __builtin_fmaf(rx, si, px[k]);                                                         \
        py[k] = __builtin_fmaf(ry, si, py[k]);                                                         \
        pz[k] = __builtin_fmaf(rz, si, pz[k]);                                                         \
        gx = __builtin_fmaf(-rx, cj, gx);                                                              \
        gy = __builtin_fmaf(-ry, cj, gy);                                                              \
        gz = __builtin_fmaf(-rz, cj, gz);                                                              \
    }                                                                                                  \
    {                                                                                                  \
        const float rx = sx - xi[KB], ry = sy - yi[KB], rz = sz - zi[KB];                              \
        const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2))); \
        const float si = pair_scale<NEWTON>(d2, UNI ? 1.0f : sm);                                      \
        px[KB] = __builtin_fmaf(rx, si, px[KB]);                                                       \
        py[KB] = __builtin_fmaf(ry, si, py[KB]);                                                       \
        pz[KB] = __builtin_fmaf(rz, si, pz[KB]);                                                       \
    }                                                                                                  \
    sx = wave_ror1<K>(sx); sy = wave_ror1<K>(sy); sz = wave_ror1<K>(sz);                               \
    if constexpr (!UNI)                                                                                \
        sm = wave_ror1<K>(sm);                                                                         \
    if constexpr (KB > 0) {                                                                            \
        gx = wave_ror1<K>(gx); gy = wave_ror1<K>(gy); gz = wave_ror1<K>(gz);                           \
    }

// Own source block KB of the wave's block (K <= 4): target slots k < KB meet
// it two-sided (the pairs of blocks k and KB, each unordered pair once), slot
// KB one-sided (the block with itself: each ordered pair once, the self pair
// adds r = 0); slots k > KB take these sources when their own block is the
// source.  The reaction sums come home to slot KB after 64 rotations.  UNI as
// in sym_block.  SPLIT: a chunk of `steps` rotation steps that ends `rot`
// lanes short of home (lane l then holds the reactions of lane l - rot's
// body); the sums are moved home by a lane shuffle.
template <int K, bool NEWTON, bool COMP, bool UNI, int KB, bool SPLIT = false>
__device__ __forceinline__ void sym_own(float4 s, const float (&xi)[K], const float (&yi)[K],
                                        const float (&zi)[K], const float (&mi)[K], float m, float eps2,
                                        float (&ax)[K], float (&ay)[K], float (&az)[K], float (&cx)[K],
                                        float (&cy)[K], float (&cz)[K], int steps = 64, int rot = 0,
                                        int lane = 0)
{
    float sx = s.x, sy = s.y, sz = s.z, sm = s.w;
    float px[KB + 1], py[KB + 1], pz[KB + 1];
#pragma unroll
    for (int k = 0; k <= KB; ++k)
        px[k] = py[k] = pz[k] = 0.f;
    float gx = 0.f, gy = 0.f, gz = 0.f;
    if constexpr (SPLIT) {
        for (int t = 0; t < steps; t += 4)  // in unrolled groups of 4, as in sym_block
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) {
                SYM_OWN_STEP();
            }
    } else {
#pragma unroll 4
        for (int t = 0; t < 64; ++t) {
            SYM_OWN_STEP();
        }
    }
#pragma unroll
    for (int k = 0; k <= KB; ++k) {
        if constexpr (UNI) {
            px[k] *= m; py[k] *= m; pz[k] *= m;
        }
        block_add<COMP>(ax[k], cx[k], px[k]);
        block_add<COMP>(ay[k], cy[k], py[k]);
        block_add<COMP>(az[k], cz[k], pz[k]);
    }
    if constexpr (KB > 0) {
        if constexpr (SPLIT) {
            if (rot != 0) {  // wave-uniform
                const int from = (lane + rot) & 63;
                gx = __shfl(gx, from, 64); gy = __shfl(gy, from, 64); gz = __shfl(gz, from, 64);
            }
        }
        if constexpr (UNI) {
            gx *= m; gy *= m; gz *= m;
        }
        block_add<COMP>(ax[KB], cx[KB], gx);
        block_add<COMP>(ay[KB], cy[KB], gy);
        block_add<COMP>(az[KB], cz[KB], gz);
    }
}

// sym_own for the wave-uniform block index kb, 0 <= kb < K.
template <int K, bool NEWTON, bool COMP, bool SPLIT = false, int KB = 0>
__device__ __forceinline__ void sym_own_at(int kb, bool uni, float4 s, const float (&xi)[K], const float (&yi)[K],
                                           const float (&zi)[K], const float (&mi)[K], float m, float eps2,
                                           float (&ax)[K], float (&ay)[K], float (&az)[K], float (&cx)[K],
                                           float (&cy)[K], float (&cz)[K], int steps = 64, int rot = 0,
                                           int lane = 0)
{
    if (kb == KB) {
        if (uni)
            sym_own<K, NEWTON, COMP, true, KB, SPLIT>(s, xi, yi, zi, mi, m, eps2, ax, ay, az, cx, cy, cz, steps,
                                                      rot, lane);
        else
            sym_own<K, NEWTON, COMP, false, KB, SPLIT>(s, xi, yi, zi, mi, m, eps2, ax, ay, az, cx, cy, cz, steps,
                                                       rot, lane);
    } else if constexpr (KB + 1 < K) {
        sym_own_at<K, NEWTON, COMP, SPLIT, KB + 1>(kb, uni, s, xi, yi, zi, mi, m, eps2, ax, ay, az, cx, cy, cz,
                                                   steps, rot, lane);
    }
}

// The work of wave r of wave block w of one system: src = the system's
// padded bodies (NW wave blocks of 64K); ipart: [R][n_pad][3] target-side
// sums; jpart: [NW][NW/2][64K][3] reactions, or the rows of wave blocks
// jbase, jbase + 1, ... only (a chunk of a step).  ipart may also hold a
// range of bodies only: body b of row r is record r * istride + b - ibody0
// (istride < 0: n_pad, the whole system).  COMP: the target-side block
// partials are Kahan-added (the compensated paths); otherwise plain FP32 sums.
// SPLIT: every sub-unit is cut into rs rotation chunks (rs divides 64); the
// reactions then take rs records per (wave block, unit): jpart is
// [NW][NW/2][rs][64K][3].
template <int K, bool NEWTON, bool COMP, bool SPLIT = false>
__device__ __forceinline__ void sym_wave(const float4* __restrict__ src, int NW, int R, int w, int r,
                                         float eps2, float* __restrict__ ipart,
                                         float* __restrict__ jpart, int lane, int jbase = 0,
                                         long long istride = -1, int ibody0 = 0, int rs = 1)
{
    const int half = NW / 2;
    const int n_pad = NW * 64 * K;
    float xi[K], yi[K], zi[K], mi[K], ax[K], ay[K], az[K], cx[K], cy[K], cz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float4 p = src[w * 64 * K + k * 64 + lane];
        xi[k] = p.x; yi[k] = p.y; zi[k] = p.z; mi[k] = p.w;
        ax[k] = ay[k] = az[k] = cx[k] = cy[k] = cz[k] = 0.f;
    }
    // one mass for every target of the wave?
    const float m0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(mi[0])));
    bool tuni = true;
#pragma unroll
    for (int k = 0; k < K; ++k)
        tuni = tuni && mi[k] == m0;
    // sub-unit q = (unit u, source block kb): u = 0 is the wave block's own
    // sources, u = 1 + d the wave block d + 1 places round
    const int Q = (half + 1) * K;
    const int QS = SPLIT ? Q * rs : Q;  // (sub-unit, chunk) pairs, chunk fastest
    const int steps = SPLIT ? 64 / rs : 64;
    const int q0 = (int)((long long)QS * r / R), q1 = (int)((long long)QS * (r + 1) / R);
    for (int qs = q0; qs < q1; ++qs) {
        int q = qs, c = 0;
        if constexpr (SPLIT) {
            q = qs / rs;
            c = qs - q * rs;
        }
        const int u = q / K, kb = q - u * K;
        // chunk c runs rotation steps [t0, t0 + steps): lane l starts with
        // source (l - t0) mod 64 and ends holding source (l - t0 - steps) mod 64
        const int t0 = c * steps;
        const int sl = SPLIT ? ((lane - t0) & 63) : lane;
        const int el = SPLIT ? ((lane - t0 - steps) & 63) : lane;
        if constexpr (K <= 4) {
            if (u == 0) {
                // own block kb: two-sided against the lower target slots
                const float4 s = src[w * 64 * K + kb * 64 + sl];
                const bool uni = __ballot(tuni && s.w == m0) == ~0ull;
                sym_own_at<K, NEWTON, COMP, SPLIT>(kb, uni, s, xi, yi, zi, mi, m0, eps2, ax, ay, az, cx, cy, cz,
                                                   steps, (t0 + steps) & 63, lane);
                continue;
            }
        }
        if constexpr (K > 4) {
            if (u == 0) {
                // own block kb (K = 8, where own blocks are a small share and
                // eight variants of sym_own would crowd the instruction cache):
                // target side only (each ordered pair once, self pair gives 0)
                static_assert(!SPLIT, "rotation chunks need K <= 4");
                const float4 s = src[w * 64 * K + kb * 64 + lane];
                float sx = s.x, sy = s.y, sz = s.z, sm = s.w;
                float px[K], py[K], pz[K];
#pragma unroll
                for (int k = 0; k < K; ++k)
                    px[k] = py[k] = pz[k] = 0.f;
#pragma unroll 4
                for (int t = 0; t < 64; ++t) {
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const float rx = sx - xi[k], ry = sy - yi[k], rz = sz - zi[k];
                        const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
                        const float si = pair_scale<NEWTON>(d2, sm);
                        px[k] = __builtin_fmaf(rx, si, px[k]);
                        py[k] = __builtin_fmaf(ry, si, py[k]);
                        pz[k] = __builtin_fmaf(rz, si, pz[k]);
                    }
                    sx = wave_ror1<K>(sx); sy = wave_ror1<K>(sy); sz = wave_ror1<K>(sz); sm = wave_ror1<K>(sm);
                }
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    block_add<COMP>(ax[k], cx[k], px[k]);
                    block_add<COMP>(ay[k], cy[k], py[k]);
                    block_add<COMP>(az[k], cz[k], pz[k]);
                }
                continue;
            }
        }
        const int d = u - 1;
        if (2 * (d + 1) == NW && w >= half)
            continue;  // the block half way round: taken by the lower half
        int v = w + 1 + d;
        if (v >= NW)
            v -= NW;
        const float4 s = src[v * 64 * K + kb * 64 + sl];
        float gx, gy, gz;
        if (__ballot(tuni && s.w == m0) == ~0ull)
            sym_block<K, NEWTON, COMP, true, SPLIT>(s, xi, yi, zi, mi, m0, eps2, ax, ay, az, cx, cy, cz, gx, gy, gz,
                                                    steps);
        else
            sym_block<K, NEWTON, COMP, false, SPLIT>(s, xi, yi, zi, mi, m0, eps2, ax, ay, az, cx, cy, cz, gx, gy,
                                                     gz, steps);
        const size_t o = ((((size_t)(w - jbase) * half + d) * (SPLIT ? rs : 1) + c) * (64 * K) + kb * 64 + el) * 3;
        jpart[o + 0] = gx;
        jpart[o + 1] = gy;
        jpart[o + 2] = gz;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const size_t o = ((size_t)r * (size_t)(istride < 0 ? n_pad : istride) + (w * 64 * K - ibody0) + k * 64 +
                          lane) * 3;
        ipart[o + 0] = ax[k] - cx[k];
        ipart[o + 1] = ay[k] - cy[k];
        ipart[o + 2] = az[k] - cz[k];
    }
}

// One system: wave gw = w * R + r.  SPLIT: rs rotation chunks per sub-unit.
template <int K, bool NEWTON, bool COMP, bool SPLIT>
__global__ __launch_bounds__(256) void sym_force_kernel(const float4* __restrict__ src, int NW, int R,
                                                        float eps2, float* __restrict__ ipart,
                                                        float* __restrict__ jpart, int rs)
{
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (gw >= NW * R)
        return;  // whole waves only
    sym_wave<K, NEWTON, COMP, SPLIT>(src, NW, R, gw / R, gw % R, eps2, ipart, jpart, threadIdx.x & 63, 0, -1, 0,
                                     rs);
}

// Batched systems: each wave runs items (system, wave block, split), a system
// laid out as for one system.
template <int K>
__global__ __launch_bounds__(256) void batch_sym_kernel(BatchSymGeom g, int nwaves, const float4* __restrict__ snap,
                                                        float eps2, float* __restrict__ ipart,
                                                        float* __restrict__ jpart)
{
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (gw >= nwaves)
        return;  // whole waves only
    const int lane = threadIdx.x & 63;
    int pos = 0;
    for (int it = batch_next_item(g, gw, pos, lane); it >= 0; it = batch_next_item(g, gw, pos, lane)) {
        const int4 item = g.items[it];  // (system, wave block, split, splits)
        const int s = item.x;
        sym_wave<K, false, true>(snap + g.off[s], g.nw[s], item.w, item.y, item.z, eps2, ipart + g.ipo[s],
                           jpart + g.jpo[s], lane);
    }
}

// Zeroes the dynamic schedule's work counter before a launch (a kernel, not a
// memset, so the step can sit in a captured graph).
__global__ void batch_counter_reset_kernel(int* counter)
{
    if (threadIdx.x == 0)
        *counter = 0;
}

// The epilogues take (hi, c) with value hi - c; split an FP64 sum exactly.
template <class Epi>
__device__ __forceinline__ void sym_epilogue(const Epi& epi, int i, const double (&s)[3])
{
    const float h0 = (float)s[0], h1 = (float)s[1], h2 = (float)s[2];
    epi(i, h0, h1, h2, (float)((double)h0 - s[0]), (float)((double)h1 - s[1]), (float)((double)h2 - s[2]));
}

// Plain FP32 sums go to the epilogue as they are (no compensation term).
template <class Epi>
__device__ __forceinline__ void sym_epilogue(const Epi& epi, int i, const float (&s)[3])
{
    epi(i, s[0], s[1], s[2], 0.f, 0.f, 0.f);
}

// Body b's partials summed by T lanes (sym_reduce.h) in FP64 (Acc = double,
// the compensated paths) or plain FP32 (Acc = float), then the epilogue.
template <int T, typename Acc, class Epi>
__global__ __launch_bounds__(256) void sym_reduce_kernel(const float* __restrict__ ipart,
                                                         const float* __restrict__ jpart, int n, int NW,
                                                         int R, int K, int rsl, Epi epi)
{
    int b, t;
    sym_reduce_lane<T>(b, t);
    Acc s[3];
    sym_split_sum<T>((const Part3*)ipart, (const Part3*)jpart, b < n ? b : n - 1, NW, R, K, t, s, rsl);
    if (b < n && t == 0)
        sym_epilogue(epi, b, s);
}

// Batched systems: body b of the concatenated padded layout (systems start on
// multiples of kBatchSymAlign, so block b / kBatchSymAlign names the system).
template <int T>
__global__ __launch_bounds__(256) void batch_sym_reduce_kernel(BatchSymGeom g, int total, int K,
                                                               const float* __restrict__ ipart,
                                                               const float* __restrict__ jpart, EpiMixed epi)
{
    int j, t;
    sym_reduce_lane<T>(j, t);
    const int b = g.body0 + j, bb = g.body0 + (j < total ? j : total - 1);
    const int s = g.sys_of_block[bb / kBatchSymAlign];
    const int loc = bb - g.off[s];  // padding bodies are summed too (their records exist)
    double sum[3];
    sym_split_sum<T>((const Part3*)(ipart + g.ipo[s]), (const Part3*)(jpart + g.jpo[s]), loc, g.nw[s], g.rs[s],
                     K, t, sum);
    if (j < total && loc < g.n[s] && t == 0)
        sym_epilogue(epi, b, sum);  // padding bodies keep their far-away zero-mass records
}

int sym_tile_r(const Tile& t, int n_pad)
{
    (void)n_pad;
    return t.s;
}

size_t sym_ipart_floats(int n_pad, int K, int R) { (void)K; return (size_t)R * n_pad * 3; }
size_t sym_jpart_floats(int n_pad, int K, int rs)
{
    const int NW = n_pad / (64 * K);
    return (size_t)NW * (NW / 2) * rs * 64 * K * 3;
}

static int log2_exact(int v)
{
    int l = 0;
    while ((1 << l) < v)
        ++l;
    return (1 << l) == v ? l : -1;
}

template <int K, bool COMP>
static void launch_sym_pairs(bool newton, int rs, dim3 grid, const float4* src, int NW, int R, float eps2,
                             float* ipart, float* jpart, hipStream_t st)
{
    if constexpr (K <= 4) {
        if (rs > 1) {
            if (newton)
                hipLaunchKernelGGL((sym_force_kernel<K, true, COMP, true>), grid, dim3(256), 0, st, src, NW, R, eps2,
                                   ipart, jpart, rs);
            else
                hipLaunchKernelGGL((sym_force_kernel<K, false, COMP, true>), grid, dim3(256), 0, st, src, NW, R, eps2,
                                   ipart, jpart, rs);
            return;
        }
    }
    if (newton)
        hipLaunchKernelGGL((sym_force_kernel<K, true, COMP, false>), grid, dim3(256), 0, st, src, NW, R, eps2,
                           ipart, jpart, 1);
    else
        hipLaunchKernelGGL((sym_force_kernel<K, false, COMP, false>), grid, dim3(256), 0, st, src, NW, R, eps2,
                           ipart, jpart, 1);
}

template <typename Acc, class Epi>
static void launch_sym_reduce(int T, const float* ipart, const float* jpart, int n, int NW, int R, int K, int rsl,
                              const Epi& epi, hipStream_t st)
{
    const dim3 rgrid((unsigned)(((long long)n * T + 255) / 256));
#define SYM_REDUCE(TT) \
    hipLaunchKernelGGL((sym_reduce_kernel<TT, Acc, Epi>), rgrid, dim3(256), 0, st, ipart, jpart, n, NW, R, K, rsl, epi)
    switch (T) {
    case 1: SYM_REDUCE(1); break;
    case 2: SYM_REDUCE(2); break;
    case 4: SYM_REDUCE(4); break;
    case 8: SYM_REDUCE(8); break;
    default: SYM_REDUCE(16); break;
    }
#undef SYM_REDUCE
}

// comp: Kahan-added block partials and an FP64 reduction (the MIXED step and
// the field of a MIXED handle); otherwise plain FP32 sums throughout
// (PREC_FP32).
template <class Epi>
hipError_t launch_sym(const Tile& t, bool newton, bool comp, const float4* src, int n, int n_pad, float eps2,
                      float* ipart, float* jpart, const Epi& epi, hipStream_t st)
{
    const int K = t.k, R = sym_tile_r(t, n_pad), rs = t.rs, rsl = log2_exact(rs);
    if (n_pad % (128 * K) != 0 || !ipart || !jpart || rsl < 0 || rs > 16 || (rs > 1 && K > 4))
        return hipErrorInvalidValue;
    const int NW = n_pad / (64 * K);
    const int waves = NW * R;
    const dim3 grid((waves + 3) / 4);
    if (K == 2)
        comp ? launch_sym_pairs<2, true>(newton, rs, grid, src, NW, R, eps2, ipart, jpart, st)
             : launch_sym_pairs<2, false>(newton, rs, grid, src, NW, R, eps2, ipart, jpart, st);
    else if (K == 4)
        comp ? launch_sym_pairs<4, true>(newton, rs, grid, src, NW, R, eps2, ipart, jpart, st)
             : launch_sym_pairs<4, false>(newton, rs, grid, src, NW, R, eps2, ipart, jpart, st);
    else if (K == 8)
        comp ? launch_sym_pairs<8, true>(newton, rs, grid, src, NW, R, eps2, ipart, jpart, st)
             : launch_sym_pairs<8, false>(newton, rs, grid, src, NW, R, eps2, ipart, jpart, st);
    else
        return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    const int T = sym_reduce_lanes(n);
    if (comp)
        launch_sym_reduce<double>(T, ipart, jpart, n, NW, R, K, rsl, epi, st);
    else
        launch_sym_reduce<float>(T, ipart, jpart, n, NW, R, K, rsl, epi, st);
    return hipGetLastError();
}

// ---- one share of a system sharded over targets by wave blocks -----------------
// Share g of G owns wave blocks [w0, w1) (the first NW % G shares one block
// more) and evaluates, for each of them, the same sub-units as sym_wave does
// in a full step: its own blocks and the half ring after it.  Its target-side
// partials go to ipart (R_share rows: the whole system's row layout, or the
// share's own contiguous region of the buffer); its reaction partials (for bodies of the wave blocks after each of its blocks,
// mostly owned by other shares) are summed per destination body into react3,
// which a sharded run adds up over the shares (a reduce-scatter); the share
// then adds its bodies' target-side partials to the summed reactions and
// integrates them.  On one device, running every share and summing the react3
// arrays gives the full step's sums in another FP64 order.

// jbase: the first wave block whose reaction rows jpart holds (0 for a
// whole-system buffer, w0 for a buffer of the share's rows only).
// ipart: rows of istride bodies starting at body ibody0 (see sym_wave).
template <int K, bool NEWTON>
__global__ __launch_bounds__(256) void sym_share_kernel(const float4* __restrict__ src, int NW, int R, int w0,
                                                        int nwaves, float eps2, float* __restrict__ ipart,
                                                        float* __restrict__ jpart, int jbase, long long istride,
                                                        int ibody0)
{
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (gw >= nwaves)
        return;  // whole waves only
    sym_wave<K, NEWTON, true>(src, NW, R, w0 + gw / R, gw % R, eps2, ipart, jpart, threadIdx.x & 63, jbase,
                              istride, ibody0);
}

// react3[b] (FP64, bodies b < n) = the reactions on body b from the wave
// blocks [w0, w1), in increasing block order; 0 where none of them reaches b.
// With accumulate, the previous react3[b] plus that sum.
__global__ __launch_bounds__(256) void sym_react_kernel(const float* __restrict__ jpart, int n, int NW, int K,
                                                        int w0, int w1, int jbase, bool accumulate,
                                                        double* __restrict__ react3)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n)
        return;
    const int half = NW / 2, B = 64 * K;
    const int v = b / B, loc = b - v * B;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int w = w0; w < w1; ++w) {
        int d = v - 1 - w;
        if (d < 0)
            d += NW;
        if (d >= half || (2 * (d + 1) == NW && w >= half))
            continue;  // v is not in w's half ring (or w's upper half skipped it)
        const float* p = jpart + (((size_t)(w - jbase) * half + d) * B + loc) * 3;
        s0 += p[0];
        s1 += p[1];
        s2 += p[2];
    }
    double* o = react3 + 3 * (size_t)b;
    if (accumulate) {
        o[0] += s0;
        o[1] += s1;
        o[2] += s2;
    } else {
        o[0] = s0;
        o[1] = s1;
        o[2] = s2;
    }
}

// Bodies [start, start + count): their R target-side partials (rows of
// istride bodies from body ibody0) summed in FP64 by T lanes per body (lane t
// takes rows t, t + T, ...; a fixed butterfly combines the lanes, as in
// sym_reduce.h), plus react_sum3[j] (the reactions summed over the shares),
// then the epilogue.
template <int T, class Epi>
__global__ __launch_bounds__(256) void sym_share_finish_kernel(const float* __restrict__ ipart, long long istride,
                                                               int ibody0, int R, int start, int count,
                                                               const double* __restrict__ react_sum3, Epi epi)
{
    int j, t;
    sym_reduce_lane<T>(j, t);
    const int b = start + (j < count ? j : count - 1);
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll 4
    for (int r = t; r < R; r += T)
        part_add(s, ((const Part3*)ipart)[(size_t)r * (size_t)istride + (b - ibody0)]);
#pragma unroll
    for (int off = 32; off >= 64 / T; off >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            s[c] += __shfl_xor(s[c], off, 64);
    if (j < count && t == 0) {
        s[0] += react_sum3[3 * (size_t)j + 0];
        s[1] += react_sum3[3 * (size_t)j + 1];
        s[2] += react_sum3[3 * (size_t)j + 2];
        sym_epilogue(epi, b, s);
    }
}

void sym_share_blocks(int NW, int shares, int share, int* w0, int* w1)
{
    const int per = NW / shares, extra = NW % shares;
    *w0 = share < extra ? share * (per + 1) : share * per + extra;
    *w1 = *w0 + (share < extra ? per + 1 : per);
}

int sym_share_splits(int nw_share, int R, int K, int NW)
{
    // about 32768 waves (ten rounds of the chip's 256 CUs x 4 SIMDs x 3 waves
    // of this kernel), every wave with the same number u of the Q sub-units
    // of its wave block (the last one of a block fewer): u = Q nw / 32768,
    // R_share = ceil(Q / u)
    (void)R;
    const long long Q = (long long)K * (NW / 2 + 1);
    long long u = Q * nw_share / 32768;
    if (u < 1)
        u = 1;
    return (int)((Q + u - 1) / u);
}

// The target-side partials of share `share` in a buffer of Rs x n_pad
// records: the whole system's rows (compact = false: body b of row r at
// r * n_pad + b) or the share's own region (compact: Rs rows of the share's
// bodies, contiguous, at record Rs * w0 * 64K; the shares' regions tile the
// buffer).
static void share_ipart_layout(int n_pad, int K, int Rs, int w0, int w1, bool compact, size_t* base,
                               long long* istride, int* ibody0)
{
    const int B = 64 * K;
    *base = compact ? (size_t)Rs * w0 * B * 3 : 0;
    *istride = compact ? (long long)(w1 - w0) * B : n_pad;
    *ibody0 = compact ? w0 * B : 0;
}

hipError_t launch_sym_share_pairs(const Tile& t, bool newton, const float4* src, int n, int n_pad, float eps2,
                                  int shares, int share, int Rs, float* ipart, bool compact_ipart, float* jpart,
                                  bool jpart_share_rows, double* react3, bool accumulate, hipStream_t st)
{
    const int K = t.k;
    if (n_pad % (128 * K) != 0 || !ipart || !jpart)
        return hipErrorInvalidValue;
    const int NW = n_pad / (64 * K);
    int w0, w1;
    sym_share_blocks(NW, shares, share, &w0, &w1);
    const int nwaves = (w1 - w0) * Rs;
    const int jbase = jpart_share_rows ? w0 : 0;
    size_t ibase;
    long long istride;
    int ibody0;
    share_ipart_layout(n_pad, K, Rs, w0, w1, compact_ipart, &ibase, &istride, &ibody0);
    if (nwaves > 0) {
        const dim3 grid((nwaves + 3) / 4), block(256);
#define SHARE_LAUNCH(KK)                                                                                  \
    if (newton)                                                                                           \
        hipLaunchKernelGGL((sym_share_kernel<KK, true>), grid, block, 0, st, src, NW, Rs, w0, nwaves, eps2, \
                           ipart + ibase, jpart, jbase, istride, ibody0);                                 \
    else                                                                                                  \
        hipLaunchKernelGGL((sym_share_kernel<KK, false>), grid, block, 0, st, src, NW, Rs, w0, nwaves, eps2, \
                           ipart + ibase, jpart, jbase, istride, ibody0);
        if (K == 2) {
            SHARE_LAUNCH(2)
        } else if (K == 4) {
            SHARE_LAUNCH(4)
        } else if (K == 8) {
            SHARE_LAUNCH(8)
        } else {
            return hipErrorInvalidValue;
        }
#undef SHARE_LAUNCH
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    hipLaunchKernelGGL(sym_react_kernel, dim3((n + 255) / 256), dim3(256), 0, st, jpart, n, NW, K, w0, w1, jbase,
                       accumulate, react3);
    return hipGetLastError();
}

template <class Epi>
static hipError_t launch_sym_finish(const float* ipart, long long istride, int ibody0, int Rs, int start, int count,
                                    const double* react_sum3, const Epi& epi, hipStream_t st)
{
    const int T = sym_reduce_lanes(count);
    const dim3 grid((unsigned)(((long long)count * T + 255) / 256)), block(256);
#define FINISH(TT) hipLaunchKernelGGL((sym_share_finish_kernel<TT, Epi>), grid, block, 0, st, ipart, istride, ibody0, \
                                      Rs, start, count, react_sum3, epi)
    switch (T) {
    case 1: FINISH(1); break;
    case 2: FINISH(2); break;
    case 4: FINISH(4); break;
    case 8: FINISH(8); break;
    default: FINISH(16); break;
    }
#undef FINISH
    return hipGetLastError();
}

// A whole step of the symmetric tile evaluated in `chunks` chunks of wave
// blocks, one after the other: each chunk's pair work (target-side partials
// into ipart, R_share rows; its reactions into jchunk, a buffer of one chunk's
// rows) and the reactions summed per body into react3, accumulated over the
// chunks in order; then every body's target-side partials plus react3 and
// the epilogue.  The same pair work as one launch of the symmetric kernel
// with partial-sum memory for one chunk's reactions.
template <class Epi>
hipError_t launch_sym_chunked(const Tile& t, bool newton, const float4* src, int n, int n_pad, float eps2,
                              int chunks, int Rs, float* ipart, float* jchunk, double* react3, const Epi& epi,
                              hipStream_t st)
{
    for (int g = 0; g < chunks; ++g) {
        hipError_t e = launch_sym_share_pairs(t, newton, src, n, n_pad, eps2, chunks, g, Rs, ipart, false, jchunk,
                                              true, react3, g > 0, st);
        if (e != hipSuccess)
            return e;
    }
    return launch_sym_finish(ipart, n_pad, 0, Rs, 0, n, react3, epi, st);
}

template hipError_t launch_sym_chunked<EpiMixed>(const Tile&, bool, const float4*, int, int, float, int, int, float*,
                                                 float*, double*, const EpiMixed&, hipStream_t);
template hipError_t launch_sym_chunked<EpiAccel>(const Tile&, bool, const float4*, int, int, float, int, int, float*,
                                                 float*, double*, const EpiAccel&, hipStream_t);

hipError_t launch_sym_share_finish_mixed(const StepArgs& a, int shares, int share, int Rs, int start, int count,
                                         const float* ipart, const double* react_sum3, float4* snap_new,
                                         double4* pos, double4* vel, hipStream_t st)
{
    const Tile& t = tile_get(a.tile);
    const int NW = a.n_pad / (64 * t.k);
    int w0, w1;
    sym_share_blocks(NW, shares, share, &w0, &w1);
    size_t ibase;
    long long istride;
    int ibody0;
    share_ipart_layout(a.n_pad, t.k, Rs, w0, w1, true, &ibase, &istride, &ibody0);
    EpiMixed epi{pos, vel, snap_new, a.dt, a.damping, a.kick};
    return launch_sym_finish(ipart + ibase, istride, ibody0, Rs, start, count, react_sum3, epi, st);
}

hipError_t launch_batch_counter_reset(int* counter, hipStream_t st)
{
    hipLaunchKernelGGL(batch_counter_reset_kernel, dim3(1), dim3(64), 0, st, counter);
    return hipGetLastError();
}

hipError_t launch_batch_sym_step(const BatchSymGeom& g, int nwaves, int total, const float4* snap_old,
                                 float4* snap_new, double4* pos, double4* vel, float eps2, double dt,
                                 double damping, int kick, float* ipart, float* jpart, hipStream_t st)
{
    hipError_t e;
    if (g.counter && (e = launch_batch_counter_reset(g.counter, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL((batch_sym_kernel<kBatchSymK>), dim3((nwaves + 3) / 4), dim3(256), 0, st, g, nwaves,
                       snap_old, eps2, ipart, jpart);
    e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    EpiMixed epi{pos, vel, snap_new, dt, damping, kick};
    const int T = sym_reduce_lanes(total);
    const dim3 rgrid((unsigned)(((long long)total * T + 255) / 256));
    switch (T) {
    case 1: hipLaunchKernelGGL(batch_sym_reduce_kernel<1>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, epi); break;
    case 2: hipLaunchKernelGGL(batch_sym_reduce_kernel<2>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, epi); break;
    case 4: hipLaunchKernelGGL(batch_sym_reduce_kernel<4>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, epi); break;
    case 8: hipLaunchKernelGGL(batch_sym_reduce_kernel<8>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, epi); break;
    default: hipLaunchKernelGGL(batch_sym_reduce_kernel<16>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, epi); break;
    }
    return hipGetLastError();
}

template hipError_t launch_sym<EpiFP32>(const Tile&, bool, bool, const float4*, int, int, float, float*, float*,
                                        const EpiFP32&, hipStream_t);
template hipError_t launch_sym<EpiMixed>(const Tile&, bool, bool, const float4*, int, int, float, float*, float*,
                                         const EpiMixed&, hipStream_t);
template hipError_t launch_sym<EpiAccel>(const Tile&, bool, bool, const float4*, int, int, float, float*, float*,
                                         const EpiAccel&, hipStream_t);

}  // namespace utcnb
