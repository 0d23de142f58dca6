// ref.hip -- the reference's own arithmetic on gfx950, bit for bit.
//
// One thread per body sums every source in index order 0..n-1 with exactly
// the operations of benchapps/nbody/seq/bodysystem.h:70-96 (sqrt_T, then a
// DOUBLE division 1.0/sqrt even for T=float, then inv*inv*inv), and then
// applies the damped kick-drift of benchapps/nbody/seq/bodysystem.cc:135-181.
// This file is compiled with -ffp-contract=off so that no multiply-add is
// fused; sqrtf and the FP64 division are IEEE correctly rounded on this
// target, so every rounding matches the reference's CPU path.
//
// Source bodies are uniform across the wave and arrive through scalar loads.
// For eps2 >= 2^-96 the FP32 kernels take ref_inv_dist (ref_inv.h): the same
// correctly rounded sqrtf and division, without the generic rescaling; for
// eps2 >= 2^-600 the FP64 kernels take ref_inv_dist64 on every group of four
// sources whose d2 stay below 2^600 (the same operations as the generic FP64
// sqrt and division there).
#include "internal.h"
#include "ref_inv.h"

namespace utcnb {

#pragma clang fp contract(off)

template <typename T, bool FAST>
__device__ __forceinline__ T inv_dist(T d2);

template <>
__device__ __forceinline__ float inv_dist<float, false>(float d2)
{
    // (float)(1.0 / (double)sqrtf(d2)) -- the literal 1.0 makes the division FP64.
    return (float)(1.0 / (double)sqrtf(d2));
}

template <>
__device__ __forceinline__ float inv_dist<float, true>(float d2)
{
    return ref_inv_dist(d2);  // d2 >= eps2 >= kRefInvMin
}

template <>
__device__ __forceinline__ double inv_dist<double, false>(double d2)
{
    return 1.0 / sqrt(d2);
}

template <>
__device__ __forceinline__ double inv_dist<double, true>(double d2)
{
    // d2 >= eps2 >= kRefInv64Min; beyond kRefInv64Max (or NaN) the generic path
    if (d2 <= kRefInv64Max)
        return ref_inv_dist64(d2);
    return 1.0 / sqrt(d2);
}

// One source j in the reference's order of operations: the geometry (r and
// d2 = ((rx rx + ry ry) + rz rz) + eps2), then the sums with inv = 1/sqrt(d2).
template <typename T, typename T4>
__device__ __forceinline__ T ref_geom(const T4& pj, const T4& pi, T eps2, T& r0, T& r1, T& r2)
{
    r0 = pj.x - pi.x;
    r1 = pj.y - pi.y;
    r2 = pj.z - pi.z;
    T d2 = r0 * r0 + r1 * r1 + r2 * r2;
    d2 += eps2;
    return d2;
}

template <typename T>
__device__ __forceinline__ void ref_accum(T mj, T inv, T r0, T r1, T r2, T& a0, T& a1, T& a2)
{
    T inv3 = inv * inv * inv;
    T s = mj * inv3;
    a0 += r0 * s;
    a1 += r1 * s;
    a2 += r2 * s;
}

template <typename T, typename T4, bool FAST>
__device__ __forceinline__ void ref_pair(const T4& pj, const T4& pi, T eps2, T& a0, T& a1, T& a2)
{
    T r0, r1, r2;
    const T d2 = ref_geom(pj, pi, eps2, r0, r1, r2);
    ref_accum(pj.w, inv_dist<T, FAST>(d2), r0, r1, r2, a0, a1, a2);
}

// G FP64 sources with one wave-uniform choice: ref_inv_dist64 when every
// lane's G d2 lie within its range (always, short of coordinates 1e90
// apart), else the generic sqrt and division.  Sums in j order as ref_pair.
template <int G>
__device__ __forceinline__ void ref_group64(const double4 (&q)[G], const double4& pi, double eps2,
                                            double& a0, double& a1, double& a2)
{
    double r[G][3], d2[G], inv[G];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        d2[k] = ref_geom(q[k], pi, eps2, r[k][0], r[k][1], r[k][2]);
        ok = ok && d2[k] <= kRefInv64Max;
    }
    if (__ballot(!ok) == 0) {
#pragma unroll
        for (int k = 0; k < G; ++k)
            inv[k] = ref_inv_dist64(d2[k]);
    } else {
#pragma unroll
        for (int k = 0; k < G; ++k)
            inv[k] = 1.0 / sqrt(d2[k]);
    }
#pragma unroll
    for (int k = 0; k < G; ++k)
        ref_accum(q[k].w, inv[k], r[k][0], r[k][1], r[k][2], a0, a1, a2);
}

// Sources in index order, sixteen per trip for FP32 (eight for FP64, whose
// inverse distance takes one uniform choice per group): their loads (one
// s_load_dwordx16 per four FP32 or two FP64 bodies) and inverse distances
// are independent, so they overlap -- with one thread per body that
// independence is the only latency hiding at small N -- while the sums still
// take j = 0, 1, 2, ... in turn (no reassociation in this file).  On MI355X
// sixteen FP32 sources ran 11% faster than four at N = 262144 and 28% at
// N = 16384..65536 (thirty-two spill SGPRs); eight FP64 sources 5% faster
// than four at 262144 and 20% at 16384..65536 (two VGPRs spill to scratch;
// sixteen spill SGPRs and ran slower).
template <typename T, typename T4, bool FAST>
__device__ __forceinline__ void ref_field(const T4* __restrict__ pos, int n, T eps2, const T4& pi,
                                          T& a0, T& a1, T& a2)
{
    a0 = 0; a1 = 0; a2 = 0;
    int j = 0;
    if constexpr (sizeof(T) == 4) {
        for (; j + 16 <= n; j += 16) {
            T4 q[16];
#pragma unroll
            for (int k = 0; k < 16; ++k)
                q[k] = pos[j + k];
#pragma unroll
            for (int k = 0; k < 16; ++k)
                ref_pair<T, T4, FAST>(q[k], pi, eps2, a0, a1, a2);
        }
    } else if constexpr (FAST) {
        for (; j + 8 <= n; j += 8) {
            T4 q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                q[k] = pos[j + k];
            ref_group64<8>(q, pi, eps2, a0, a1, a2);
        }
    }
    for (; j + 4 <= n; j += 4) {
        const T4 q0 = pos[j], q1 = pos[j + 1], q2 = pos[j + 2], q3 = pos[j + 3];
        if constexpr (FAST && sizeof(T) == 8) {
            const T4 q[4] = {q0, q1, q2, q3};
            ref_group64<4>(q, pi, eps2, a0, a1, a2);
        } else {
            ref_pair<T, T4, FAST>(q0, pi, eps2, a0, a1, a2);
            ref_pair<T, T4, FAST>(q1, pi, eps2, a0, a1, a2);
            ref_pair<T, T4, FAST>(q2, pi, eps2, a0, a1, a2);
            ref_pair<T, T4, FAST>(q3, pi, eps2, a0, a1, a2);
        }
    }
    for (; j < n; ++j)
        ref_pair<T, T4, FAST>(pos[j], pi, eps2, a0, a1, a2);
}

// Fused force + integrate; the twin of the reference's
// _integrateNBodySystem_kernel<T> (benchapps/nbody/cuda/bodysystem_kernel.cu:87-138)
// and of its multi-GPU form that takes a slice of targets
// (benchapps/nbody/utc-snode-mgpu/mgpu/bodysystem_kernel.cu:88-140): targets
// i0 + t for t < cnt against all n sources of pos_old; the new position and
// velocity of target i0 + t are pos_new[t] and vel[t] (the caller offsets
// full-size arrays by i0).
template <typename T, typename T4, bool FAST>
__global__ void ref_step_kernel(const T4* __restrict__ pos_old, T4* __restrict__ pos_new,
                                T4* __restrict__ vel, int n, int i0, int cnt, T dt, T eps2, T damping)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt)
        return;
    T4 p = pos_old[i0 + t];
    T a0, a1, a2;
    ref_field<T, T4, FAST>(pos_old, n, eps2, p, a0, a1, a2);
    T w = p.w;  // the reference's "invMass" (holds m_i), seq/bodysystem.cc:149
    T4 v = vel[t];
    v.x += (a0 * w) * dt;
    v.y += (a1 * w) * dt;
    v.z += (a2 * w) * dt;
    v.x *= damping;
    v.y *= damping;
    v.z *= damping;
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    pos_new[t] = p;
    vel[t] = v;
}

// Field of targets i0 + t, t < cnt, into acc3[3t .. 3t+2].
template <typename T, typename T4, bool FAST>
__global__ void ref_accel_kernel(const T4* __restrict__ pos, int n, int i0, int cnt, T eps2,
                                 double* __restrict__ acc3)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt)
        return;
    T a0, a1, a2;
    ref_field<T, T4, FAST>(pos, n, eps2, pos[i0 + t], a0, a1, a2);
    acc3[3 * (size_t)t + 0] = (double)a0;
    acc3[3 * (size_t)t + 1] = (double)a1;
    acc3[3 * (size_t)t + 2] = (double)a2;
}

hipError_t launch_ref_step_slice_f32(const float4* pos_old, float4* pos_new, float4* vel, int n, int i0, int cnt,
                                     float dt, float eps2, float damping, int block, hipStream_t st)
{
    if (block <= 0)
        block = 256;
    const dim3 grid((unsigned)((cnt + block - 1) / block));
    if (eps2 >= kRefInvMin && eps2 <= 3.0e38f)
        hipLaunchKernelGGL((ref_step_kernel<float, float4, true>), grid, dim3(block), 0, st, pos_old, pos_new,
                           vel, n, i0, cnt, dt, eps2, damping);
    else
        hipLaunchKernelGGL((ref_step_kernel<float, float4, false>), grid, dim3(block), 0, st, pos_old, pos_new,
                           vel, n, i0, cnt, dt, eps2, damping);
    return hipGetLastError();
}

hipError_t launch_ref_step_slice_f64(const double4* pos_old, double4* pos_new, double4* vel, int n, int i0,
                                     int cnt, double dt, double eps2, double damping, int block, hipStream_t st)
{
    if (block <= 0)
        block = 256;
    const dim3 grid((unsigned)((cnt + block - 1) / block));
    if (eps2 >= kRefInv64Min)
        hipLaunchKernelGGL((ref_step_kernel<double, double4, true>), grid, dim3(block), 0, st, pos_old, pos_new,
                           vel, n, i0, cnt, dt, eps2, damping);
    else
        hipLaunchKernelGGL((ref_step_kernel<double, double4, false>), grid, dim3(block), 0, st, pos_old, pos_new,
                           vel, n, i0, cnt, dt, eps2, damping);
    return hipGetLastError();
}

hipError_t launch_ref_step_f32(const float4* pos_old, float4* pos_new, float4* vel, int n,
                               float dt, float eps2, float damping, int block, hipStream_t st)
{
    return launch_ref_step_slice_f32(pos_old, pos_new, vel, n, 0, n, dt, eps2, damping, block, st);
}

hipError_t launch_ref_step_f64(const double4* pos_old, double4* pos_new, double4* vel, int n,
                               double dt, double eps2, double damping, int block,
                               hipStream_t st)
{
    return launch_ref_step_slice_f64(pos_old, pos_new, vel, n, 0, n, dt, eps2, damping, block, st);
}

hipError_t launch_ref_accel_slice_f32(const float4* src, int n, int i0, int cnt, float eps2, double* acc3,
                                      hipStream_t st)
{
    const dim3 grid((unsigned)((cnt + 255) / 256));
    if (eps2 >= kRefInvMin && eps2 <= 3.0e38f)
        hipLaunchKernelGGL((ref_accel_kernel<float, float4, true>), grid, dim3(256), 0, st, src, n, i0, cnt, eps2,
                           acc3);
    else
        hipLaunchKernelGGL((ref_accel_kernel<float, float4, false>), grid, dim3(256), 0, st, src, n, i0, cnt, eps2,
                           acc3);
    return hipGetLastError();
}

hipError_t launch_ref_accel_slice_f64(const double4* src, int n, int i0, int cnt, double eps2, double* acc3,
                                      hipStream_t st)
{
    const dim3 grid((unsigned)((cnt + 255) / 256));
    if (eps2 >= kRefInv64Min)
        hipLaunchKernelGGL((ref_accel_kernel<double, double4, true>), grid, dim3(256), 0, st, src, n, i0, cnt,
                           eps2, acc3);
    else
        hipLaunchKernelGGL((ref_accel_kernel<double, double4, false>), grid, dim3(256), 0, st, src, n, i0, cnt,
                           eps2, acc3);
    return hipGetLastError();
}

hipError_t launch_ref_accel_f32(const float4* src, int n, float eps2, double* acc3,
                                hipStream_t st)
{
    return launch_ref_accel_slice_f32(src, n, 0, n, eps2, acc3, st);
}

hipError_t launch_ref_accel_f64(const double4* src, int n, double eps2, double* acc3,
                                hipStream_t st)
{
    return launch_ref_accel_slice_f64(src, n, 0, n, eps2, acc3, st);
}

}  // namespace utcnb
