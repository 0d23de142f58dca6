// sym_reduce.h -- per-body sums of the symmetric kernels' partials (device
// code only), shared by the force (sym.hip) and VJP (sym_vjp.hip) reductions.
//
// Body b of a system with NW wave blocks of 64K bodies has Q = R + RS NW/2
// partial records: q < R is wave r = q's target-side partial (ipart, R rows
// of n_pad records), q = R + RS d + c the reaction from rotation chunk c of
// the wave block d + 1 places below (jpart, RS = 2^rsl chunks per sub-unit;
// RS = 1 for the kernels without a rotation split), absent for the block half
// way round when the upper half skipped it.  T lanes share a body: lane t adds records t, t + T, t + 2T, ...
// in that order (in FP64 on the compensated paths, in FP32 on the plain one), then the T partial sums are combined by a butterfly
// of __shfl_xor over fixed lane offsets.  The order depends only on (T, R,
// NW), so results are bitwise repeatable; with T = 1 it is the plain
// sequential order.  Lane layout in a wave: lane = t * (64/T) + j for the
// wave's j-th body, so the 64/T lanes of one t read 64/T consecutive records.
#pragma once

#include <hip/hip_runtime.h>

namespace utcnb {

struct Part3 {
    float x, y, z;
};

template <typename Acc>
__device__ __forceinline__ void part_add(Acc (&s)[3], const Part3& p)
{
    s[0] += p.x; s[1] += p.y; s[2] += p.z;
}
__device__ __forceinline__ void part_add(double (&s)[4], const float4& p)
{
    s[0] += p.x; s[1] += p.y; s[2] += p.z; s[3] += p.w;
}

// Lanes per body: the fewest (1..16) that give the reduction about 2^19
// threads, enough waves to keep HBM requests in flight at small N.
inline int sym_reduce_lanes(long long bodies)
{
    int T = 1;
    while (T < 16 && bodies * T < (1LL << 19))
        T *= 2;
    return T;
}

// (body, t) of this thread for T lanes per body; bodies are numbered from the
// first body of the launch.
template <int T>
__device__ __forceinline__ void sym_reduce_lane(int& body, int& t)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = g & 63;
    t = lane / (64 / T);
    body = (g >> 6) * (64 / T) + lane % (64 / T);
}

// All 64 lanes of the wave must call this (it shuffles); b must be a valid
// body of the system for every lane.
template <int T, typename P, typename Acc, int C>
__device__ __forceinline__ void sym_split_sum(const P* __restrict__ ipart, const P* __restrict__ jpart, int b,
                                              int NW, int R, int K, int t, Acc (&s)[C], int rsl = 0)
{
    const int n_pad = NW * 64 * K;
    const int half = NW / 2, v = b / (64 * K), loc = b - v * (64 * K);
#pragma unroll
    for (int c = 0; c < C; ++c)
        s[c] = Acc(0);
    const int Q = R + (half << rsl);
#pragma unroll 4
    for (int q = t; q < Q; q += T) {
        const P* p;
        if (q < R) {
            p = ipart + (size_t)q * n_pad + b;
        } else {
            const int dc = q - R, d = dc >> rsl;
            int w = v - 1 - d;
            if (w < 0)
                w += NW;
            if (2 * (d + 1) == NW && w >= half)
                continue;
            p = jpart + ((((size_t)w * half + d) << rsl) + (dc - (d << rsl))) * (64 * K) + loc;
        }
        part_add(s, *p);
    }
#pragma unroll
    for (int off = 32; off >= 64 / T; off >>= 1)
#pragma unroll
        for (int c = 0; c < C; ++c)
            s[c] += __shfl_xor(s[c], off, 64);
}

}  // namespace utcnb
