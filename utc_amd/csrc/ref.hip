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
#include "internal.h"

namespace utcnb {

#pragma clang fp contract(off)

template <typename T>
__device__ __forceinline__ T inv_dist(T d2);

template <>
__device__ __forceinline__ float inv_dist<float>(float d2)
{
    // (float)(1.0 / (double)sqrtf(d2)) -- the literal 1.0 makes the division FP64.
    return (float)(1.0 / (double)sqrtf(d2));
}

template <>
__device__ __forceinline__ double inv_dist<double>(double d2)
{
    return 1.0 / sqrt(d2);
}

template <typename T, typename T4>
__device__ __forceinline__ void ref_field(const T4* __restrict__ pos, int n, T eps2, const T4& pi,
                                          T& a0, T& a1, T& a2)
{
    a0 = 0; a1 = 0; a2 = 0;
    for (int j = 0; j < n; ++j) {
        const T4 pj = pos[j];
        T r0 = pj.x - pi.x;
        T r1 = pj.y - pi.y;
        T r2 = pj.z - pi.z;
        T d2 = r0 * r0 + r1 * r1 + r2 * r2;
        d2 += eps2;
        T inv = inv_dist<T>(d2);
        T inv3 = inv * inv * inv;
        T s = pj.w * inv3;
        a0 += r0 * s;
        a1 += r1 * s;
        a2 += r2 * s;
    }
}

// Fused force + integrate; the twin of the reference's
// _integrateNBodySystem_kernel<T> (benchapps/nbody/cuda/bodysystem_kernel.cu:87-138).
template <typename T, typename T4>
__global__ void ref_step_kernel(const T4* __restrict__ pos_old, T4* __restrict__ pos_new,
                                T4* __restrict__ vel, int n, T dt, T eps2, T damping)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    T4 p = pos_old[i];
    T a0, a1, a2;
    ref_field<T, T4>(pos_old, n, eps2, p, a0, a1, a2);
    T w = p.w;  // the reference's "invMass" (holds m_i), seq/bodysystem.cc:149
    T4 v = vel[i];
    v.x += (a0 * w) * dt;
    v.y += (a1 * w) * dt;
    v.z += (a2 * w) * dt;
    v.x *= damping;
    v.y *= damping;
    v.z *= damping;
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    pos_new[i] = p;
    vel[i] = v;
}

template <typename T, typename T4>
__global__ void ref_accel_kernel(const T4* __restrict__ pos, int n, T eps2,
                                 double* __restrict__ acc3)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    T a0, a1, a2;
    ref_field<T, T4>(pos, n, eps2, pos[i], a0, a1, a2);
    acc3[3 * (size_t)i + 0] = (double)a0;
    acc3[3 * (size_t)i + 1] = (double)a1;
    acc3[3 * (size_t)i + 2] = (double)a2;
}

hipError_t launch_ref_step_f32(const float4* pos_old, float4* pos_new, float4* vel, int n,
                               float dt, float eps2, float damping, int block, hipStream_t st)
{
    if (block <= 0)
        block = 256;
    hipLaunchKernelGGL((ref_step_kernel<float, float4>), dim3((n + block - 1) / block),
                       dim3(block), 0, st, pos_old, pos_new, vel, n, dt, eps2, damping);
    return hipGetLastError();
}

hipError_t launch_ref_step_f64(const double4* pos_old, double4* pos_new, double4* vel, int n,
                               double dt, double eps2, double damping, int block,
                               hipStream_t st)
{
    if (block <= 0)
        block = 256;
    hipLaunchKernelGGL((ref_step_kernel<double, double4>), dim3((n + block - 1) / block),
                       dim3(block), 0, st, pos_old, pos_new, vel, n, dt, eps2, damping);
    return hipGetLastError();
}

hipError_t launch_ref_accel_f32(const float4* src, int n, float eps2, double* acc3,
                                hipStream_t st)
{
    hipLaunchKernelGGL((ref_accel_kernel<float, float4>), dim3((n + 255) / 256), dim3(256), 0,
                       st, src, n, eps2, acc3);
    return hipGetLastError();
}

hipError_t launch_ref_accel_f64(const double4* src, int n, double eps2, double* acc3,
                                hipStream_t st)
{
    hipLaunchKernelGGL((ref_accel_kernel<double, double4>), dim3((n + 255) / 256), dim3(256),
                       0, st, src, n, eps2, acc3);
    return hipGetLastError();
}

}  // namespace utcnb
