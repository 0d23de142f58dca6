// force_small.h -- one-kernel steps for small systems (device code; included
// by force_tmpl.h).
//
// Below about 8192 bodies a step is bound by latency, not by issue rate: the
// symmetric kernels need a second kernel to add their partial sums, and the
// one-pass kernels give every wave 64 targets, so a small system has too few
// waves to fill the chip and each wave waits on its scalar source loads.
// Here a workgroup of W waves owns TW = TL * A targets and all sources:
//  * the workgroup copies every source body (n_pad <= CAP) into LDS once;
//  * lane l = t + TL * g of wave w holds targets t, t + TL, ..., t + (A-1) TL
//    of the workgroup and sums sources split, split + SPL, split + 2 SPL, ...
//    with split = w * G + g, G = 64 / TL source groups per wave and SPL =
//    G * W source splits per workgroup: the TL lanes of a group read the same
//    LDS record (a broadcast), the G groups consecutive ones;
//  * the G lanes of a target meet in a butterfly of lane shuffles, the W
//    waves in LDS, in a fixed order (bitwise repeatable), and TW threads run
//    the integrator epilogue: one launch per step, nothing in HBM between.
// COMP: sums of up to kSmallBlock sources are Kahan-added per lane (as the
// 64-source partials of the other kernels) and the lanes and waves are
// combined in FP64; otherwise plain FP32 throughout (the plain FP32 path).
// Pair arithmetic as in stream_sources: 12 VALU slots per ordered pair plus
// one LDS read per source shared by the lane's A targets.
#pragma once

#include <type_traits>

namespace utcnb {

constexpr int kSmallBlock = 32;  // sources per plain partial of a lane (COMP)

template <int A, int TL, int W, int CAP, bool COMP, bool NEWTON, class Epi>
__global__ __launch_bounds__(64 * W) void small_force_kernel(const float4* __restrict__ src, int i0, int nt,
                                                             int n_pad, float eps2, Epi epi)
{
    constexpr int G = 64 / TL, TW = TL * A, SPL = G * W;
    using Acc = typename std::conditional<COMP, double, float>::type;
    __shared__ float4 sh[CAP];
    __shared__ Acc red[W][TW][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane % TL, g = lane / TL;
    const int tgt0 = blockIdx.x * TW;
    // the epilogue's state loads, issued before the pair work (thread j < TW
    // updates target tgt0 + j)
    typename Epi::State pre{};
    if (threadIdx.x < TW && tgt0 + (int)threadIdx.x < nt)
        pre = epi.load(i0 + tgt0 + threadIdx.x);

    float px[A], py[A], pz[A], hx[A], hy[A], hz[A], cx[A], cy[A], cz[A];
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int tt = tgt0 + a * TL + t;
        const float4 p = src[i0 + (tt < nt ? tt : nt - 1)];  // clamped lanes compute but never store
        px[a] = p.x; py[a] = p.y; pz[a] = p.z;
        hx[a] = hy[a] = hz[a] = cx[a] = cy[a] = cz[a] = 0.f;
    }
    for (int k = threadIdx.x; k < n_pad; k += 64 * W)
        sh[k] = src[k];
    __syncthreads();

    const int split = wave * G + g;
    const int iters = n_pad / SPL;  // n_pad is a multiple of 1024 and SPL divides 1024
    for (int b0 = 0; b0 < iters; b0 += kSmallBlock) {
        const int b1 = b0 + kSmallBlock < iters ? b0 + kSmallBlock : iters;
        float ax[A], ay[A], az[A];
#pragma unroll
        for (int a = 0; a < A; ++a) {
            if constexpr (COMP) {
                ax[a] = ay[a] = az[a] = 0.f;
            } else {
                ax[a] = hx[a]; ay[a] = hy[a]; az[a] = hz[a];
            }
        }
#pragma unroll 8
        for (int i = b0; i < b1; ++i) {
            const float4 b = sh[split + i * SPL];
#pragma unroll
            for (int a = 0; a < A; ++a) {
                const float rx = b.x - px[a], ry = b.y - py[a], rz = b.z - pz[a];
                const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
                const float s = pair_scale<NEWTON>(d2, b.w);
                ax[a] = __builtin_fmaf(rx, s, ax[a]);
                ay[a] = __builtin_fmaf(ry, s, ay[a]);
                az[a] = __builtin_fmaf(rz, s, az[a]);
            }
        }
#pragma unroll
        for (int a = 0; a < A; ++a) {
            if constexpr (COMP) {
                kahan_add(hx[a], cx[a], ax[a]);
                kahan_add(hy[a], cy[a], ay[a]);
                kahan_add(hz[a], cz[a], az[a]);
            } else {
                hx[a] = ax[a]; hy[a] = ay[a]; hz[a] = az[a];
            }
        }
    }

    // lanes t, t + TL, ..., t + 64 - TL hold the same targets: butterfly
#pragma unroll
    for (int a = 0; a < A; ++a) {
        Acc v[3];
        if constexpr (COMP) {  // Kahan keeps the running error as -c: the sum is hi - c
            v[0] = (double)hx[a] - (double)cx[a];
            v[1] = (double)hy[a] - (double)cy[a];
            v[2] = (double)hz[a] - (double)cz[a];
        } else {
            v[0] = hx[a]; v[1] = hy[a]; v[2] = hz[a];
        }
#pragma unroll
        for (int off = TL; off < 64; off <<= 1)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                v[c] += __shfl_xor(v[c], off, 64);
        if (g == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                red[wave][a * TL + t][c] = v[c];
        }
    }
    __syncthreads();
    if (threadIdx.x < TW) {
        const int j = threadIdx.x;
        Acc s[3] = {red[0][j][0], red[0][j][1], red[0][j][2]};
#pragma unroll
        for (int w = 1; w < W; ++w)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                s[c] += red[w][j][c];
        const int tt = tgt0 + j;
        if (tt < nt) {
            if constexpr (COMP) {  // split the FP64 sum exactly into (hi, c) with hi - c = s
                const float h0 = (float)s[0], h1 = (float)s[1], h2 = (float)s[2];
                epi.store(i0 + tt, pre, h0, h1, h2, (float)((double)h0 - s[0]), (float)((double)h1 - s[1]),
                          (float)((double)h2 - s[2]));
            } else {
                epi.store(i0 + tt, pre, s[0], s[1], s[2], 0.f, 0.f, 0.f);
            }
        }
    }
}

// Targets [i0, i0 + nt) of a system whose n_pad sources fit the tile's LDS
// capacity CAP.
template <int A, int TL, int W, int CAP, bool COMP, bool NEWTON, class Epi>
hipError_t launch_small_tile(const float4* src, int i0, int nt, int n_pad, float eps2, const Epi& epi,
                             hipStream_t st)
{
    static_assert(64 % TL == 0 && 1024 % ((64 / TL) * W) == 0, "source splits must divide kPadJ");
    if (n_pad > CAP || n_pad % kPadJ != 0)
        return hipErrorInvalidValue;
    constexpr int TW = TL * A;
    const int blocks = (nt + TW - 1) / TW;
    hipLaunchKernelGGL((small_force_kernel<A, TL, W, CAP, COMP, NEWTON, Epi>), dim3(blocks), dim3(64 * W), 0, st,
                       src, i0, nt, n_pad, eps2, epi);
    return hipGetLastError();
}

}  // namespace utcnb
