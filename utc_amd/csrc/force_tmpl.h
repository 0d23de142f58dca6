// force_tmpl.h -- all-pairs softened-gravity kernels for gfx950 (CDNA4).
//
// Mapping.  A wavefront owns K target bodies per lane (64*K targets) and
// streams a contiguous range of source bodies.  The source index is uniform
// across the wave, so the compiler loads source bodies with scalar loads
// (s_load_dwordx16 = four float4 bodies) into SGPRs, and every VALU op takes
// its source operand straight from an SGPR: no LDS traffic and no broadcast
// instructions in the inner loop.  S waves of a workgroup split the source
// range of one target tile; their partial sums meet in LDS at the end.
//
// Arithmetic per pair (the reference's pair, benchapps/nbody/seq/bodysystem.h:70-96,
// rewritten for the VALU): r = pj - pi (3 sub), d2 = r.r + eps2 (3 fma),
// y = v_rsq_f32(d2) (transcendental; it co-issues with the FMA pipe),
// s = mj*y*y*y (3 mul), acc += r*s (3 fma): 12 VALU issue slots per pair.
// Measured on MI355X this runs at ~99% of the non-packed FP32 VALU issue
// rate; replacing y^3 by rsq(d2)*rcp(d2) (11 slots, two transcendentals) is
// slower because the second transcendental does not co-issue.  The NEWTON
// variant applies one Newton-Raphson step to y, carried into the cube: with
// e = 1 - d2*y^2 the refined y1 = y (1 + e/2) has y1^3 = y^3 (1 + 3e/2) +
// O(e^2) y^3, and |e| ~ 2^-22 leaves the e^2 terms far below FP32 rounding,
// so y1^3 = fma(y^3 * e, 1.5, y^3): three slots more than y^3 alone (15 per
// pair).  Refining y first and cubing it (e, y/2, fma, then the cube) costs
// four, and y*(1.5 - 0.5*d2*y*y) five.
//
// Compensation.  Each target keeps (hi, c) per component: 64 sources are
// summed plainly into a partial, which is added to hi with Kahan's
// compensated summation (Kahan 1965).  Cost: 12 ops per 64 pairs.
#pragma once

#include "internal.h"

namespace utcnb {

__device__ __forceinline__ float rsq_f32(float x) { return __builtin_amdgcn_rsqf(x); }

template <bool NEWTON>
__device__ __forceinline__ float pair_scale(float d2, float m)
{
    if constexpr (NEWTON) {
        const float y = rsq_f32(d2);
        const float y2 = y * y;
        const float e = __builtin_fmaf(-d2, y2, 1.0f);
        const float y3 = y2 * y;
        return m * __builtin_fmaf(y3 * e, 1.5f, y3);
    } else {
        float y = rsq_f32(d2);
        return m * y * y * y;
    }
}

// Kahan step: add partial p into (s, c).
__device__ __forceinline__ void kahan_add(float& s, float& c, float p)
{
    float y = p - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Stream sources [j0, j1) (multiple of kChunk) against K targets per lane.
template <int K, bool COMP, bool NEWTON>
__device__ __forceinline__ void stream_sources(const float4* __restrict__ src, int j0, int j1,
                                               float eps2, const float (&px)[K],
                                               const float (&py)[K], const float (&pz)[K],
                                               float (&hx)[K], float (&hy)[K], float (&hz)[K],
                                               float (&cx)[K], float (&cy)[K], float (&cz)[K])
{
    for (int jc = j0; jc < j1; jc += kChunk) {
        float ax[K], ay[K], az[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if constexpr (COMP) {
                ax[k] = 0.f; ay[k] = 0.f; az[k] = 0.f;
            } else {
                ax[k] = hx[k]; ay[k] = hy[k]; az[k] = hz[k];
            }
        }
#pragma unroll 4
        for (int j = jc; j < jc + kChunk; ++j) {
            const float4 b = src[j];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                float rx = b.x - px[k];
                float ry = b.y - py[k];
                float rz = b.z - pz[k];
                float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
                float s = pair_scale<NEWTON>(d2, b.w);
                ax[k] = __builtin_fmaf(rx, s, ax[k]);
                ay[k] = __builtin_fmaf(ry, s, ay[k]);
                az[k] = __builtin_fmaf(rz, s, az[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if constexpr (COMP) {
                kahan_add(hx[k], cx[k], ax[k]);
                kahan_add(hy[k], cy[k], ay[k]);
                kahan_add(hz[k], cz[k], az[k]);
            } else {
                hx[k] = ax[k]; hy[k] = ay[k]; hz[k] = az[k];
            }
        }
    }
}

// ---- epilogues ---------------------------------------------------------------
// Called once per target i < n with the summed field a = hi + c' where c' is
// the compensation term (zero for plain sums).  load(i) reads the state the
// update needs and store(i, state, ...) applies it, so a kernel can issue the
// loads early; operator() does both.

// Plain FP32 state, the reference's damped kick-drift
// (benchapps/nbody/seq/bodysystem.cc:161-172) with kick factor pos.w or 1.
struct EpiFP32 {
    const float4* pos_old;
    float4* pos_new;
    float4* vel;
    float dt, damping;
    int kick;
    struct State {
        float4 p, v;
    };
    __device__ __forceinline__ State load(int i) const { return {pos_old[i], vel[i]}; }
    __device__ __forceinline__ void store(int i, State st, float hx, float hy, float hz, float, float,
                                          float) const
    {
        float4 p = st.p;
        float4 v = st.v;
        float w = kick == UTCNB_KICK_REFERENCE ? p.w : 1.0f;
        v.x = (v.x + (hx * w) * dt) * damping;
        v.y = (v.y + (hy * w) * dt) * damping;
        v.z = (v.z + (hz * w) * dt) * damping;
        p.x += v.x * dt;
        p.y += v.y * dt;
        p.z += v.z * dt;
        pos_new[i] = p;
        vel[i] = v;
    }
    __device__ __forceinline__ void operator()(int i, float hx, float hy, float hz, float cx, float cy,
                                               float cz) const
    {
        store(i, load(i), hx, hy, hz, cx, cy, cz);
    }
};

// FP64 state updated in place, new FP32 snapshot written for the next step.
struct EpiMixed {
    double4* pos;
    double4* vel;
    float4* snap_new;
    double dt, damping;
    int kick;
    struct State {
        double4 p, v;
    };
    __device__ __forceinline__ State load(int i) const { return {pos[i], vel[i]}; }
    __device__ __forceinline__ void store(int i, State st, float hx, float hy, float hz, float cx, float cy,
                                          float cz) const
    {
        // Kahan keeps the running error as -c; the sum is hi - c.
        double ax = (double)hx - (double)cx;
        double ay = (double)hy - (double)cy;
        double az = (double)hz - (double)cz;
        double4 p = st.p;
        double4 v = st.v;
        double w = kick == UTCNB_KICK_REFERENCE ? p.w : 1.0;
        v.x = (v.x + (ax * w) * dt) * damping;
        v.y = (v.y + (ay * w) * dt) * damping;
        v.z = (v.z + (az * w) * dt) * damping;
        p.x += v.x * dt;
        p.y += v.y * dt;
        p.z += v.z * dt;
        pos[i] = p;
        vel[i] = v;
        snap_new[i] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
    }
    __device__ __forceinline__ void operator()(int i, float hx, float hy, float hz, float cx, float cy,
                                               float cz) const
    {
        store(i, load(i), hx, hy, hz, cx, cy, cz);
    }
};

// Field only, as FP64 [3n]; body i goes to row i - off (off = the first
// target of a slice, 0 for all bodies).
struct EpiAccel {
    double* acc3;
    int off = 0;
    struct State {};
    __device__ __forceinline__ State load(int) const { return {}; }
    __device__ __forceinline__ void store(int i, State, float hx, float hy, float hz, float cx, float cy,
                                          float cz) const
    {
        const size_t r = 3 * (size_t)(i - off);
        acc3[r + 0] = (double)hx - (double)cx;
        acc3[r + 1] = (double)hy - (double)cy;
        acc3[r + 2] = (double)hz - (double)cz;
    }
    __device__ __forceinline__ void operator()(int i, float hx, float hy, float hz, float cx, float cy,
                                               float cz) const
    {
        store(i, load(i), hx, hy, hz, cx, cy, cz);
    }
};

}  // namespace utcnb

#include "force_small.h"  // uses pair_scale and kahan_add above

namespace utcnb {

// ---- the kernel --------------------------------------------------------------
// W waves per workgroup; S of them split the sources of one target tile, so a
// workgroup holds W/S target tiles of 64*K bodies.  `src` holds n_pad bodies
// (multiple of kPadJ); the targets are bodies [i0, i0 + nt) (all n bodies, or
// one slice of them, against every source).
template <int K, int S, int W, bool COMP, bool NEWTON, class Epi>
__global__ __launch_bounds__(64 * W) void force_kernel(const float4* __restrict__ src, int i0, int nt,
                                                       int n_pad, float eps2, Epi epi)
{
    constexpr int T = W / S;  // target tiles per workgroup
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tile = blockIdx.x * T + wave / S;
    const int part = wave % S;

    float px[K], py[K], pz[K];
    float hx[K], hy[K], hz[K], cx[K], cy[K], cz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        int t = tile * 64 * K + k * 64 + lane;
        int il = i0 + (t < nt ? t : nt - 1);  // clamp: lanes past the last target compute but never store
        float4 p = src[il];
        px[k] = p.x; py[k] = p.y; pz[k] = p.z;
        hx[k] = hy[k] = hz[k] = 0.f;
        cx[k] = cy[k] = cz[k] = 0.f;
    }
    const int len = n_pad / S;
    stream_sources<K, COMP, NEWTON>(src, part * len, part * len + len, eps2, px, py, pz, hx, hy,
                                    hz, cx, cy, cz);

    if constexpr (S > 1) {
        // Partials of the S waves meet in LDS; wave part 0 combines them in
        // part order (deterministic).
        __shared__ float red[W][K][6][64];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            red[wave][k][0][lane] = hx[k];
            red[wave][k][1][lane] = hy[k];
            red[wave][k][2][lane] = hz[k];
            red[wave][k][3][lane] = cx[k];
            red[wave][k][4][lane] = cy[k];
            red[wave][k][5][lane] = cz[k];
        }
        __syncthreads();
        if (part != 0)
            return;
#pragma unroll
        for (int q = 1; q < S; ++q) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if constexpr (COMP) {
                    // fold the other wave's (hi, c): hi - c is its value
                    kahan_add(hx[k], cx[k], red[wave + q][k][0][lane] - red[wave + q][k][3][lane]);
                    kahan_add(hy[k], cy[k], red[wave + q][k][1][lane] - red[wave + q][k][4][lane]);
                    kahan_add(hz[k], cz[k], red[wave + q][k][2][lane] - red[wave + q][k][5][lane]);
                } else {
                    hx[k] += red[wave + q][k][0][lane];
                    hy[k] += red[wave + q][k][1][lane];
                    hz[k] += red[wave + q][k][2][lane];
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        int t = tile * 64 * K + k * 64 + lane;
        if (t < nt)
            epi(i0 + t, hx[k], hy[k], hz[k], cx[k], cy[k], cz[k]);
    }
}

template <int K, int S, bool COMP, bool NEWTON, class Epi>
hipError_t launch_force_tile(const float4* src, int i0, int nt, int n_pad, float eps2, const Epi& epi,
                             hipStream_t st)
{
    constexpr int W = S < 4 ? 4 : S;
    constexpr int T = W / S;
    const int tiles = (nt + 64 * K - 1) / (64 * K);
    const int blocks = (tiles + T - 1) / T;
    hipLaunchKernelGGL((force_kernel<K, S, W, COMP, NEWTON, Epi>), dim3(blocks), dim3(64 * W), 0,
                       st, src, i0, nt, n_pad, eps2, epi);
    return hipGetLastError();
}

// Dispatch over the tile table (must match tile_get()); targets [i0, i0 + nt).
template <bool COMP, bool NEWTON, class Epi>
hipError_t launch_force_range(int tile, const float4* src, int i0, int nt, int n_pad, float eps2,
                              const Epi& epi, hipStream_t st)
{
    if (nt < 1)
        return hipErrorInvalidValue;
    switch (tile) {
    case 0: return launch_force_tile<1, 1, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 1: return launch_force_tile<2, 1, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 2: return launch_force_tile<4, 1, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 3: return launch_force_tile<1, 2, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 4: return launch_force_tile<2, 2, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 5: return launch_force_tile<1, 4, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 6: return launch_force_tile<2, 4, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 7: return launch_force_tile<1, 8, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 8: return launch_force_tile<2, 8, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 9: return launch_force_tile<1, 16, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 28: return launch_small_tile<1, 4, 8, 2048, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 29: return launch_small_tile<1, 8, 8, 2048, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 30: return launch_small_tile<1, 8, 16, 8192, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 31: return launch_small_tile<1, 16, 16, 4096, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 32: return launch_small_tile<1, 16, 16, 8192, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 33: return launch_small_tile<2, 16, 16, 8192, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 34: return launch_force_tile<2, 16, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 35: return launch_force_tile<4, 8, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    case 36: return launch_force_tile<4, 16, COMP, NEWTON>(src, i0, nt, n_pad, eps2, epi, st);
    default: return hipErrorInvalidValue;
    }
}

// All n bodies as targets.
template <bool COMP, bool NEWTON, class Epi>
hipError_t launch_force(int tile, const float4* src, int n, int n_pad, float eps2,
                        const Epi& epi, hipStream_t st)
{
    return launch_force_range<COMP, NEWTON>(tile, src, 0, n, n_pad, eps2, epi, st);
}

}  // namespace utcnb
