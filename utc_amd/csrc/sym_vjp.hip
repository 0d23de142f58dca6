// sym_vjp.hip -- the adjoint pair sum (VJP) evaluating every unordered pair once.
//
// With alpha the per-body kick adjoint, r = x_j - x_i, d^2 = |r|^2 + eps2,
// y = d^-1, J(r) = y^3 I - 3 y^5 r r^T and w = m_i alpha_j - m_j alpha_i, the
// pair (i, j) adds
//   xb_i += J w,            xb_j -= J w            (J is even in r, w odd in i<->j)
//   mb_i -= y^3 alpha_j.r,  mb_j += y^3 alpha_i.r
// so both bodies' terms come from one evaluation.  Written as
// J w = y^3 (w - 3 y^2 (r.w) r) with r.w = m_i (a_j.r) - m_j (a_i.r), reusing
// the two dot products the mass terms need, an unordered pair costs 36 VALU
// slots (3 sub, 3 fma for d^2, rsq, 2 mul for y^2 and y^3, 6 for the two dot
// products, 2 for r.w, 2 mul for -3 y^2 r.w, 6 for w, 3 fma for the bracket,
// 6 fma for the two position terms, 2 fma for the two mass terms) against 31
// per ordered pair in the one-pass VJP, plus 11 DPP moves per 64-source
// rotation step shared by the K targets of a lane.  When the targets of the
// wave and the sources of a block all have one mass m, w = m (alpha_j -
// alpha_i) and r.w = m (a_j.r - a_i.r): the pair works with the differences
// alone and the block's position partials are scaled by m once (31 slots, 10
// DPP moves).
//
// The mapping is the symmetric force kernel's (sym.hip): wave blocks of 64*K
// bodies, unit 0 = own blocks (source block kb two-sided against target slots
// k < kb, one-sided against slot kb), unit 1 + d = the wave block d + 1 places
// round (both sides), R waves per wave block split the K (NW/2+1) (unit,
// source block) sub-units.
// Partials go to HBM (float4 per body: x, y, z, m) and a second kernel adds
// them per body in FP64, in a fixed order, into xbar and mbar.
#include "internal.h"
#include "sym_items.h"
#include "sym_reduce.h"
#include "vjp_pair.h"

namespace utcnb {

namespace {

// Lane l receives lane l-1's value (lane 0 lane 63's).  Every lane has a
// source under wave_ror:1, so the "old" operand is never used; passing v lets
// the compiler rotate in place, without a v_mov that first zeroes a new
// register: 8% faster VJP steps at N = 262144 on MI355X than "old" = 0 (a
// same-device A/B; the force kernel measured the other way, see sym.hip).
__device__ __forceinline__ float ror1(float v)
{
    const int x = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(x, x, 0x13C /* wave_ror:1 */, 0xF, 0xF, false));
}

__device__ __forceinline__ void kahan1(float& s, float& c, float p)
{
    const float y = p - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

}  // namespace

// A block of 64 sources (one per lane: position/mass sp, alpha sa) against
// the K targets of each lane, both sides: the target-side block partials are
// Kahan-added into (g, c); h receives the source's reaction sums after 64
// rotations.  UNI: every target mass mi[k] and source mass equals m.
template <int K, bool UNI>
__device__ __forceinline__ void sym_vjp_block(float4 sp, float4 sa, const float (&xi)[K], const float (&yi)[K],
                                              const float (&zi)[K], const float (&mi)[K], const float (&bx)[K],
                                              const float (&by)[K], const float (&bz)[K], float m, float eps2,
                                              float (&g)[K][4], float (&c)[K][4], float4& h)
{
    float sx = sp.x, sy = sp.y, sz = sp.z, sm = sp.w, ax = sa.x, ay = sa.y, az = sa.z;
    float hx = 0.f, hy = 0.f, hz = 0.f, hm = 0.f;  // reaction on the source body
    float p[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k)
        p[k][0] = p[k][1] = p[k][2] = p[k][3] = 0.f;
#pragma unroll 2
    for (int t = 0; t < 64; ++t) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float rx = sx - xi[k], ry = sy - yi[k], rz = sz - zi[k];
            const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
            const float y = __builtin_amdgcn_rsqf(d2);
            const float y2 = y * y, y3 = y2 * y;
            // J w = y^3 (w - 3 y^2 (r.w) r), r.w = m_i (a_j.r) - m_j (a_i.r)
            const float arj = __builtin_fmaf(ax, rx, __builtin_fmaf(ay, ry, az * rz));
            const float ari = __builtin_fmaf(bx[k], rx, __builtin_fmaf(by[k], ry, bz[k] * rz));
            float s3, wx, wy, wz;
            if constexpr (UNI) {  // w / m and r.w / m
                s3 = (-3.f * y2) * (arj - ari);
                wx = ax - bx[k]; wy = ay - by[k]; wz = az - bz[k];
            } else {
                s3 = (-3.f * y2) * __builtin_fmaf(mi[k], arj, -(sm * ari));
                wx = w_fast(mi[k], ax, sm, bx[k]);
                wy = w_fast(mi[k], ay, sm, by[k]);
                wz = w_fast(mi[k], az, sm, bz[k]);
            }
            const float tx = __builtin_fmaf(s3, rx, wx);
            const float ty = __builtin_fmaf(s3, ry, wy);
            const float tz = __builtin_fmaf(s3, rz, wz);
            p[k][0] = __builtin_fmaf(y3, tx, p[k][0]);
            p[k][1] = __builtin_fmaf(y3, ty, p[k][1]);
            p[k][2] = __builtin_fmaf(y3, tz, p[k][2]);
            hx = __builtin_fmaf(-y3, tx, hx);
            hy = __builtin_fmaf(-y3, ty, hy);
            hz = __builtin_fmaf(-y3, tz, hz);
            p[k][3] = __builtin_fmaf(-y3, arj, p[k][3]);
            hm = __builtin_fmaf(y3, ari, hm);
        }
        sx = ror1(sx); sy = ror1(sy); sz = ror1(sz);
        if constexpr (!UNI)
            sm = ror1(sm);
        ax = ror1(ax); ay = ror1(ay); az = ror1(az);
        hx = ror1(hx); hy = ror1(hy); hz = ror1(hz); hm = ror1(hm);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if constexpr (UNI) {
            p[k][0] *= m; p[k][1] *= m; p[k][2] *= m;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            kahan1(g[k][q], c[k][q], p[k][q]);
    }
    if constexpr (UNI) {
        hx *= m; hy *= m; hz *= m;
    }
    h = make_float4(hx, hy, hz, hm);
}

// Own source block KB of the wave's block: target slots k < KB meet it
// two-sided (as sym_vjp_block; no self pairs), slot KB one-sided (the block
// with itself, self pairs included, so w is exact: vjp_pair.h); slots k > KB
// take these sources when their own block is the source.  The reaction sums
// come home to slot KB after 64 rotations.  UNI applies to the two-sided
// pairs.
template <int K, bool UNI, int KB>
__device__ __forceinline__ void sym_vjp_own(float4 sp, float4 sa, const float (&xi)[K], const float (&yi)[K],
                                            const float (&zi)[K], const float (&mi)[K], const float (&bx)[K],
                                            const float (&by)[K], const float (&bz)[K], float m, float eps2,
                                            float (&g)[K][4], float (&c)[K][4])
{
    float sx = sp.x, sy = sp.y, sz = sp.z, sm = sp.w, ax = sa.x, ay = sa.y, az = sa.z;
    float hx = 0.f, hy = 0.f, hz = 0.f, hm = 0.f;  // reaction on the source body
    float p[KB + 1][4];
#pragma unroll
    for (int k = 0; k <= KB; ++k)
        p[k][0] = p[k][1] = p[k][2] = p[k][3] = 0.f;
#pragma unroll 2
    for (int t = 0; t < 64; ++t) {
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const float rx = sx - xi[k], ry = sy - yi[k], rz = sz - zi[k];
            const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
            const float y = __builtin_amdgcn_rsqf(d2);
            const float y2 = y * y, y3 = y2 * y;
            const float arj = __builtin_fmaf(ax, rx, __builtin_fmaf(ay, ry, az * rz));
            const float ari = __builtin_fmaf(bx[k], rx, __builtin_fmaf(by[k], ry, bz[k] * rz));
            float s3, wx, wy, wz;
            if constexpr (UNI) {
                s3 = (-3.f * y2) * (arj - ari);
                wx = ax - bx[k]; wy = ay - by[k]; wz = az - bz[k];
            } else {
                s3 = (-3.f * y2) * __builtin_fmaf(mi[k], arj, -(sm * ari));
                wx = w_fast(mi[k], ax, sm, bx[k]);
                wy = w_fast(mi[k], ay, sm, by[k]);
                wz = w_fast(mi[k], az, sm, bz[k]);
            }
            const float tx = __builtin_fmaf(s3, rx, wx);
            const float ty = __builtin_fmaf(s3, ry, wy);
            const float tz = __builtin_fmaf(s3, rz, wz);
            p[k][0] = __builtin_fmaf(y3, tx, p[k][0]);
            p[k][1] = __builtin_fmaf(y3, ty, p[k][1]);
            p[k][2] = __builtin_fmaf(y3, tz, p[k][2]);
            hx = __builtin_fmaf(-y3, tx, hx);
            hy = __builtin_fmaf(-y3, ty, hy);
            hz = __builtin_fmaf(-y3, tz, hz);
            p[k][3] = __builtin_fmaf(-y3, arj, p[k][3]);
            hm = __builtin_fmaf(y3, ari, hm);
        }
        {
            const float rx = sx - xi[KB], ry = sy - yi[KB], rz = sz - zi[KB];
            const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
            const float y = __builtin_amdgcn_rsqf(d2);
            const float y2 = y * y, y3 = y2 * y, y5m3 = (-3.f * y2) * y3;
            const float wx = w_exact(mi[KB], ax, sm, bx[KB]);
            const float wy = w_exact(mi[KB], ay, sm, by[KB]);
            const float wz = w_exact(mi[KB], az, sm, bz[KB]);
            const float q = y5m3 * __builtin_fmaf(rx, wx, __builtin_fmaf(ry, wy, rz * wz));
            p[KB][0] = __builtin_fmaf(q, rx, __builtin_fmaf(y3, wx, p[KB][0]));
            p[KB][1] = __builtin_fmaf(q, ry, __builtin_fmaf(y3, wy, p[KB][1]));
            p[KB][2] = __builtin_fmaf(q, rz, __builtin_fmaf(y3, wz, p[KB][2]));
            const float arj = __builtin_fmaf(ax, rx, __builtin_fmaf(ay, ry, az * rz));
            p[KB][3] = __builtin_fmaf(-y3, arj, p[KB][3]);
        }
        sx = ror1(sx); sy = ror1(sy); sz = ror1(sz); sm = ror1(sm);
        ax = ror1(ax); ay = ror1(ay); az = ror1(az);
        if constexpr (KB > 0) {
            hx = ror1(hx); hy = ror1(hy); hz = ror1(hz); hm = ror1(hm);
        }
    }
#pragma unroll
    for (int k = 0; k <= KB; ++k) {
        if constexpr (UNI) {
            if (k < KB) {
                p[k][0] *= m; p[k][1] *= m; p[k][2] *= m;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            kahan1(g[k][q], c[k][q], p[k][q]);
    }
    if constexpr (KB > 0) {
        if constexpr (UNI) {
            hx *= m; hy *= m; hz *= m;
        }
        kahan1(g[KB][0], c[KB][0], hx);
        kahan1(g[KB][1], c[KB][1], hy);
        kahan1(g[KB][2], c[KB][2], hz);
        kahan1(g[KB][3], c[KB][3], hm);
    }
}

// sym_vjp_own for the wave-uniform block index kb, 0 <= kb < K.
template <int K, int KB = 0>
__device__ __forceinline__ void sym_vjp_own_at(int kb, bool uni, float4 sp, float4 sa, const float (&xi)[K],
                                               const float (&yi)[K], const float (&zi)[K], const float (&mi)[K],
                                               const float (&bx)[K], const float (&by)[K], const float (&bz)[K],
                                               float m, float eps2, float (&g)[K][4], float (&c)[K][4])
{
    if (kb == KB) {
        if (uni)
            sym_vjp_own<K, true, KB>(sp, sa, xi, yi, zi, mi, bx, by, bz, m, eps2, g, c);
        else
            sym_vjp_own<K, false, KB>(sp, sa, xi, yi, zi, mi, bx, by, bz, m, eps2, g, c);
    } else if constexpr (KB + 1 < K) {
        sym_vjp_own_at<K, KB + 1>(kb, uni, sp, sa, xi, yi, zi, mi, bx, by, bz, m, eps2, g, c);
    }
}

// The work of wave r of wave block w of one system.  jrec: two float4 per
// body {x,y,z,m} {ax,ay,az,0}, n_pad = NW*64*K bodies; ipart: [R][n_pad]
// float4; jpart: [NW][NW/2][64K] float4.
template <int K>
__device__ __forceinline__ void sym_vjp_wave(const float4* __restrict__ jrec, int NW, int R, int w, int r,
                                             float eps2, float4* __restrict__ ipart,
                                             float4* __restrict__ jpart, int lane)
{
    const int half = NW / 2;
    const int n_pad = NW * 64 * K;
    float xi[K], yi[K], zi[K], mi[K], bx[K], by[K], bz[K];
    float g[K][4], c[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int b = w * 64 * K + k * 64 + lane;
        const float4 p = jrec[2 * b], a = jrec[2 * b + 1];
        xi[k] = p.x; yi[k] = p.y; zi[k] = p.z; mi[k] = p.w;
        bx[k] = a.x; by[k] = a.y; bz[k] = a.z;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            g[k][q] = c[k][q] = 0.f;
    }
    // one mass m0 for every target of the wave?
    const float m0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(mi[0])));
    bool tuni = true;
#pragma unroll
    for (int k = 0; k < K; ++k)
        tuni = tuni && mi[k] == m0;
    // sub-unit su = (unit u, source block kb): u = 0 is the wave block's own
    // sources, u = 1 + d the wave block d + 1 places round
    const int Q = (half + 1) * K;
    const int q0 = (int)((long long)Q * r / R), q1 = (int)((long long)Q * (r + 1) / R);
    for (int su = q0; su < q1; ++su) {
        const int u = su / K, kb = su - u * K;
        if (u == 0) {
            // own block kb: two-sided against the lower target slots
            const int b = w * 64 * K + kb * 64 + lane;
            const float4 sp = jrec[2 * b], sa = jrec[2 * b + 1];
            const bool uni = __ballot(tuni && sp.w == m0) == ~0ull;
            sym_vjp_own_at<K>(kb, uni, sp, sa, xi, yi, zi, mi, bx, by, bz, m0, eps2, g, c);
            continue;
        }
        const int d = u - 1;
        if (2 * (d + 1) == NW && w >= half)
            continue;  // the block half way round: taken by the lower half
        int v = w + 1 + d;
        if (v >= NW)
            v -= NW;
        const int b = v * 64 * K + kb * 64 + lane;
        const float4 sp = jrec[2 * b], sa = jrec[2 * b + 1];
        float4 h;
        if (__ballot(tuni && sp.w == m0) == ~0ull)
            sym_vjp_block<K, true>(sp, sa, xi, yi, zi, mi, bx, by, bz, m0, eps2, g, c, h);
        else
            sym_vjp_block<K, false>(sp, sa, xi, yi, zi, mi, bx, by, bz, m0, eps2, g, c, h);
        jpart[((size_t)w * half + d) * (64 * K) + kb * 64 + lane] = h;
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        ipart[(size_t)r * n_pad + w * 64 * K + k * 64 + lane] =
            make_float4(g[k][0] - c[k][0], g[k][1] - c[k][1], g[k][2] - c[k][2], g[k][3] - c[k][3]);
}

// One system: wave gw = w * R + r.
template <int K>
__global__ __launch_bounds__(256) void sym_vjp_kernel(const float4* __restrict__ jrec, int NW, int R,
                                                      float eps2, float4* __restrict__ ipart,
                                                      float4* __restrict__ jpart)
{
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (gw >= NW * R)
        return;  // whole waves only
    sym_vjp_wave<K>(jrec, NW, R, gw / R, gw % R, eps2, ipart, jpart, threadIdx.x & 63);
}

// Batched systems: each wave runs items in the schedule of batch_next_item
// (sym_items.h).
template <int K>
__global__ __launch_bounds__(256) void batch_sym_vjp_kernel(BatchSymGeom g, int nwaves,
                                                            const float4* __restrict__ jrec, float eps2,
                                                            float4* __restrict__ ipart,
                                                            float4* __restrict__ jpart)
{
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (gw >= nwaves)
        return;  // whole waves only
    const int lane = threadIdx.x & 63;
    int pos = 0;
    for (int it = batch_next_item(g, gw, pos, lane); it >= 0; it = batch_next_item(g, gw, pos, lane)) {
        const int4 item = g.items[it];  // (system, wave block, split, splits)
        const int s = item.x;
        sym_vjp_wave<K>(jrec + 2 * (size_t)g.off[s], g.nw[s], item.w, item.y, item.z, eps2,
                        ipart + g.vipo[s], jpart + g.vjpo[s], lane);
    }
}

// xbar_b += the sum of body b's partials, mbar_b likewise (T lanes per body,
// sym_reduce.h).
template <int T>
__global__ __launch_bounds__(256) void sym_vjp_reduce_kernel(const float4* __restrict__ ipart,
                                                             const float4* __restrict__ jpart, int n,
                                                             int NW, int R, int K,
                                                             double* __restrict__ xbar3,
                                                             double* __restrict__ mbar)
{
    int b, t;
    sym_reduce_lane<T>(b, t);
    double s[4];
    sym_split_sum<T>(ipart, jpart, b < n ? b : n - 1, NW, R, K, t, s);
    if (b < n && t == 0) {
        xbar3[3 * (size_t)b + 0] += s[0];
        xbar3[3 * (size_t)b + 1] += s[1];
        xbar3[3 * (size_t)b + 2] += s[2];
        mbar[b] += s[3];
    }
}

// Batched systems: body b of the concatenated padded layout.
template <int T>
__global__ __launch_bounds__(256) void batch_sym_vjp_reduce_kernel(BatchSymGeom g, int total, int K,
                                                                   const float4* __restrict__ ipart,
                                                                   const float4* __restrict__ jpart,
                                                                   double* __restrict__ xbar3,
                                                                   double* __restrict__ mbar)
{
    int j, t;
    sym_reduce_lane<T>(j, t);
    const int b = g.body0 + j, bb = g.body0 + (j < total ? j : total - 1);
    const int s = g.sys_of_block[bb / kBatchSymAlign];
    const int loc = bb - g.off[s];
    double sum[4];
    sym_split_sum<T>(ipart + g.vipo[s], jpart + g.vjpo[s], loc, g.nw[s], g.rs[s], K, t, sum);
    if (j < total && loc < g.n[s] && t == 0) {
        xbar3[3 * (size_t)b + 0] += sum[0];
        xbar3[3 * (size_t)b + 1] += sum[1];
        xbar3[3 * (size_t)b + 2] += sum[2];
        mbar[b] += sum[3];
    }
}

constexpr int kSymVjpK = 4;

int sym_vjp_splits(int n_pad, int cu_count)
{
    // ~16 waves per SIMD, at least one source block per wave, at most 32
    // splits (sweeps of R on MI355X: for the adjoint at N = 16384, 50 steps,
    // checkpoint every 10, with the segment recomputes beside the VJP steps,
    // 32 is 2.6% faster than 48 and 0.6% faster than 64; at 65536 16..64 are
    // the same within noise; N = 262144 takes R = 16 either way)
    const int NW = n_pad / (64 * kSymVjpK);
    const int want = (cu_count > 0 ? cu_count : 256) * 4 * 16;
    int R = (want + NW - 1) / NW;
    const int sub = (NW / 2 + 1) * kSymVjpK;
    if (R > sub)
        R = sub;
    if (R > 32)
        R = 32;
    return R < 1 ? 1 : R;
}

size_t sym_vjp_ipart_float4s(int n_pad, int R) { return (size_t)R * n_pad; }
size_t sym_vjp_jpart_float4s(int n_pad)
{
    const int NW = n_pad / (64 * kSymVjpK);
    return (size_t)NW * (NW / 2) * 64 * kSymVjpK;
}
bool sym_vjp_applies(int n_pad) { return n_pad % (128 * kSymVjpK) == 0; }

hipError_t launch_sym_vjp(const float4* jrec, int n, int n_pad, float eps2, int R, float4* ipart,
                          float4* jpart, double* xbar3, double* mbar, hipStream_t st)
{
    if (!sym_vjp_applies(n_pad) || R < 1 || !ipart || !jpart)
        return hipErrorInvalidValue;
    const int NW = n_pad / (64 * kSymVjpK);
    const int waves = NW * R;
    hipLaunchKernelGGL((sym_vjp_kernel<kSymVjpK>), dim3((waves + 3) / 4), dim3(256), 0, st, jrec, NW, R,
                       eps2, ipart, jpart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    const int T = sym_reduce_lanes(n);
    const dim3 rgrid((unsigned)(((long long)n * T + 255) / 256));
    switch (T) {
    case 1: hipLaunchKernelGGL(sym_vjp_reduce_kernel<1>, rgrid, dim3(256), 0, st, ipart, jpart, n, NW, R, kSymVjpK, xbar3, mbar); break;
    case 2: hipLaunchKernelGGL(sym_vjp_reduce_kernel<2>, rgrid, dim3(256), 0, st, ipart, jpart, n, NW, R, kSymVjpK, xbar3, mbar); break;
    case 4: hipLaunchKernelGGL(sym_vjp_reduce_kernel<4>, rgrid, dim3(256), 0, st, ipart, jpart, n, NW, R, kSymVjpK, xbar3, mbar); break;
    case 8: hipLaunchKernelGGL(sym_vjp_reduce_kernel<8>, rgrid, dim3(256), 0, st, ipart, jpart, n, NW, R, kSymVjpK, xbar3, mbar); break;
    default: hipLaunchKernelGGL(sym_vjp_reduce_kernel<16>, rgrid, dim3(256), 0, st, ipart, jpart, n, NW, R, kSymVjpK, xbar3, mbar); break;
    }
    return hipGetLastError();
}

hipError_t launch_batch_sym_vjp(const BatchSymGeom& g, int nwaves, int total, const float4* jrec,
                                float eps2, float4* ipart, float4* jpart, double* xbar3, double* mbar,
                                hipStream_t st)
{
    static_assert(kBatchSymK == kSymVjpK, "batched and single-system symmetric VJP share K");
    hipError_t e;
    if (g.counter && (e = launch_batch_counter_reset(g.counter, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL((batch_sym_vjp_kernel<kBatchSymK>), dim3((nwaves + 3) / 4), dim3(256), 0, st, g,
                       nwaves, jrec, eps2, ipart, jpart);
    e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    const int T = sym_reduce_lanes(total);
    const dim3 rgrid((unsigned)(((long long)total * T + 255) / 256));
    switch (T) {
    case 1: hipLaunchKernelGGL(batch_sym_vjp_reduce_kernel<1>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, xbar3, mbar); break;
    case 2: hipLaunchKernelGGL(batch_sym_vjp_reduce_kernel<2>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, xbar3, mbar); break;
    case 4: hipLaunchKernelGGL(batch_sym_vjp_reduce_kernel<4>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, xbar3, mbar); break;
    case 8: hipLaunchKernelGGL(batch_sym_vjp_reduce_kernel<8>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, xbar3, mbar); break;
    default: hipLaunchKernelGGL(batch_sym_vjp_reduce_kernel<16>, rgrid, dim3(256), 0, st, g, total, kBatchSymK, ipart, jpart, xbar3, mbar); break;
    }
    return hipGetLastError();
}

}  // namespace utcnb
