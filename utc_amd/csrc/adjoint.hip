// adjoint.hip -- reverse mode through the force evaluation and the damped
// kick-drift, for the loss L = (1/n) sum_i |x_i(T) - t_i|^2.
//
// Forward step (reference form, benchapps/nbody/seq/bodysystem.cc:135-181,
// with kick factor k_i = m_i (reference) or 1 (physical)):
//   a_i = sum_j m_j K(r_ij),  K(r) = r d^-3,  d^2 = |r|^2 + eps2,  r_ij = x_j - x_i
//   u_i = D (v_i + k_i a_i dt);  x_i' = x_i + u_i dt;  v_i' = u_i
// Adjoint of one step, given (xb', vb'):
//   ub = vb' + xb' dt;  vb = D ub;  alpha_i = D k_i dt ub_i
//   mb_i += D dt ub_i . a_i                                (reference kick only)
//   xb_i = xb'_i + sum_j J(r_ij) (m_i alpha_j - m_j alpha_i),  J = d^-3 I - 3 d^-5 r r^T
//   mb_i += -sum_j alpha_j . K(r_ij)
// The pair sum (the VJP) has the same gather shape as the force: per target
// i, stream every source j.  Sources are uniform across a wave and come in
// through scalar loads of an 8-float record (x, y, z, m, ax, ay, az, 0).
// Per pair: 30 VALU ops + one v_rsq (FP32) or a correctly rounded sqrt and
// division (FP64), with Kahan-compensated 64-source chunks.
#include "internal.h"
#include "ref_inv.h"
#include "vjp_pair.h"

namespace utcnb {

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    using type = float4;
};
template <>
struct Vec4<double> {
    using type = double4;
};

template <typename T>
__device__ __forceinline__ T inv_sqrt(T d2);
template <>
__device__ __forceinline__ float inv_sqrt<float>(float d2)
{
    return __builtin_amdgcn_rsqf(d2);
}
template <>
__device__ __forceinline__ double inv_sqrt<double>(double d2)
{
    // the correctly rounded 1/sqrt; ref_inv_dist64 gives the same bits with
    // fewer operations when every lane's d2 is in [2^-600, 2^600] (ref_inv.h)
    if (__ballot(!(d2 >= kRefInv64Min && d2 <= kRefInv64Max)) == 0)
        return ref_inv_dist64(d2);
    return 1.0 / __builtin_sqrt(d2);
}

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ __forceinline__ void kahan(T& s, T& c, T p)
{
    T y = p - c;
    T t = s + y;
    c = (t - s) - y;
    s = t;
}

// One chunk of 64 sources into the partial sums p (x, y, z, m) of target i.
// EXACT: the chunk may hold j = i (vjp_pair.h).
template <bool EXACT, typename T, typename T4>
__device__ __forceinline__ void vjp_chunk(const T4* __restrict__ jrec, int jc, const T4& pi,
                                          const T4& ai, T eps2, T p[4])
{
#pragma unroll 2
    for (int j = jc; j < jc + kChunk; ++j) {
        const T4 pj = jrec[2 * j];
        const T4 aj = jrec[2 * j + 1];
        T rx = pj.x - pi.x, ry = pj.y - pi.y, rz = pj.z - pi.z;
        T d2 = fma_(rx, rx, fma_(ry, ry, fma_(rz, rz, eps2)));
        T y = inv_sqrt<T>(d2);
        T y2 = y * y;
        T y3 = y2 * y;
        T y5 = y3 * y2;
        // w = m_i alpha_j - m_j alpha_i
        T wx = w_term<EXACT>(pi.w, aj.x, pj.w, ai.x);
        T wy = w_term<EXACT>(pi.w, aj.y, pj.w, ai.y);
        T wz = w_term<EXACT>(pi.w, aj.z, pj.w, ai.z);
        T rw = fma_(rx, wx, fma_(ry, wy, rz * wz));
        T q = (T)(-3) * y5 * rw;
        p[0] = fma_(q, rx, fma_(y3, wx, p[0]));
        p[1] = fma_(q, ry, fma_(y3, wy, p[1]));
        p[2] = fma_(q, rz, fma_(y3, wz, p[2]));
        T ar = fma_(aj.x, rx, fma_(aj.y, ry, aj.z * rz));
        p[3] = fma_(-y3, ar, p[3]);
    }
}

// jrec: per body two T4 records {x,y,z,m} {ax,ay,az,0}; n_pad bodies (padding:
// zero mass, zero alpha, far away).  xbar (double4 per body: x,y,z,-) and mbar
// are accumulated in place for the first n bodies.
template <typename T, int S, int W>
__global__ __launch_bounds__(64 * W) void vjp_kernel(const typename Vec4<T>::type* __restrict__ jrec,
                                                     int n, int n_pad, T eps2,
                                                     double* __restrict__ xbar3,
                                                     double* __restrict__ mbar)
{
    using T4 = typename Vec4<T>::type;
    constexpr int TT = W / S;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tile = blockIdx.x * TT + wave / S;
    const int part = wave % S;
    const int i = tile * 64 + lane;
    const int il = i < n ? i : n - 1;
    const T4 pi = jrec[2 * il];
    const T4 ai = jrec[2 * il + 1];
    T g[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    const int len = n_pad / S;
    const int j0 = part * len;
    const int jdiag = tile * 64;  // the chunk holding j = i for every lane of the wave
    for (int jc = j0; jc < j0 + len; jc += kChunk) {
        T p[4] = {0, 0, 0, 0};
        if (jc == jdiag)
            vjp_chunk<true>(jrec, jc, pi, ai, eps2, p);
        else
            vjp_chunk<false>(jrec, jc, pi, ai, eps2, p);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            kahan(g[q], c[q], p[q]);
    }
    if constexpr (S > 1) {
        __shared__ T red[W][8][64];  // FP64, S = 16: 64 KiB
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            red[wave][q][lane] = g[q];
            red[wave][4 + q][lane] = c[q];
        }
        __syncthreads();
        if (part != 0)
            return;
        for (int o = 1; o < S; ++o)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                kahan(g[q], c[q], red[wave + o][q][lane] - red[wave + o][4 + q][lane]);
    }
    if (i < n) {
        xbar3[3 * (size_t)i + 0] += (double)g[0] - (double)c[0];
        xbar3[3 * (size_t)i + 1] += (double)g[1] - (double)c[1];
        xbar3[3 * (size_t)i + 2] += (double)g[2] - (double)c[2];
        mbar[i] += (double)g[3] - (double)c[3];
    }
}

template <typename T, int S>
static hipError_t launch_vjp_s(const void* jrec, int n, int n_pad, double eps2, double* xbar3,
                               double* mbar, hipStream_t st)
{
    constexpr int W = S < 4 ? 4 : S;
    constexpr int TT = W / S;
    const int tiles = (n + 63) / 64;
    const int blocks = (tiles + TT - 1) / TT;
    using T4 = typename Vec4<T>::type;
    hipLaunchKernelGGL((vjp_kernel<T, S, W>), dim3(blocks), dim3(64 * W), 0, st,
                       (const T4*)jrec, n, n_pad, (T)eps2, xbar3, mbar);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_vjp(const void* jrec, int n, int n_pad, double eps2, double* xbar3,
                             double* mbar, int cu_count, hipStream_t st)
{
    // as for the force kernel: split the sources until ~16 waves per SIMD,
    // keeping every split a whole number of 64-body chunks
    const int64_t waves = (n + 63) / 64;
    const int64_t want = (int64_t)(cu_count > 0 ? cu_count : 256) * 4 * 16;
    int S = waves * 4 >= want ? 4 : (waves * 8 >= want ? 8 : 16);
    while (S > 1 && n_pad % (kChunk * S) != 0)
        S /= 2;
    switch (S) {
    case 1: return launch_vjp_s<T, 1>(jrec, n, n_pad, eps2, xbar3, mbar, st);
    case 2: return launch_vjp_s<T, 2>(jrec, n, n_pad, eps2, xbar3, mbar, st);
    case 4: return launch_vjp_s<T, 4>(jrec, n, n_pad, eps2, xbar3, mbar, st);
    case 8: return launch_vjp_s<T, 8>(jrec, n, n_pad, eps2, xbar3, mbar, st);
    default: return launch_vjp_s<T, 16>(jrec, n, n_pad, eps2, xbar3, mbar, st);
    }
}

hipError_t launch_vjp_f32(const float4* jrec, int n, int n_pad, double eps2, double* xbar3,
                          double* mbar, int cu_count, hipStream_t st)
{
    return launch_vjp<float>(jrec, n, n_pad, eps2, xbar3, mbar, cu_count, st);
}

hipError_t launch_vjp_f64(const double4* jrec, int n, int n_pad, double eps2, double* xbar3,
                          double* mbar, int cu_count, hipStream_t st)
{
    return launch_vjp<double>(jrec, n, n_pad, eps2, xbar3, mbar, cu_count, st);
}

// ---- per-body kernels ------------------------------------------------------------

// Loss seed: xbar = 2 (x - t) / n, vbar = 0, mbar = 0; per-block partial sums
// of |x - t|^2 into part[blockIdx] (reduced by loss_finish).
__global__ __launch_bounds__(256) void loss_seed_kernel(const double4* __restrict__ pos,
                                                        const double* __restrict__ tgt3, int n,
                                                        double* __restrict__ xbar3,
                                                        double* __restrict__ vbar3,
                                                        double* __restrict__ mbar,
                                                        double* __restrict__ part)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        double4 p = pos[i];
        double dx = p.x - tgt3[3 * (size_t)i + 0];
        double dy = p.y - tgt3[3 * (size_t)i + 1];
        double dz = p.z - tgt3[3 * (size_t)i + 2];
        s = dx * dx + dy * dy + dz * dz;
        const double f = 2.0 / n;
        xbar3[3 * (size_t)i + 0] = f * dx;
        xbar3[3 * (size_t)i + 1] = f * dy;
        xbar3[3 * (size_t)i + 2] = f * dz;
        vbar3[3 * (size_t)i + 0] = 0.0;
        vbar3[3 * (size_t)i + 1] = 0.0;
        vbar3[3 * (size_t)i + 2] = 0.0;
        mbar[i] = 0.0;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        s += __shfl_xor(s, off, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(64) void loss_finish_kernel(const double* __restrict__ part,
                                                         int nparts, int n,
                                                         double* __restrict__ loss)
{
    double s = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 64) {
        double y = part[b] - c;
        double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    double v = s - c;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off, 64);
    if (threadIdx.x == 0)
        *loss = v / n;
}

// One step's per-body adjoint and the source records for the VJP.
// acc3 (nullable) is the field at this step's positions (reference kick only).
template <typename T4, typename E>
__global__ __launch_bounds__(256) void adjoint_pre_kernel(const double4* __restrict__ pos,
                                                          const double* __restrict__ acc3,
                                                          int n, double dt, double damping,
                                                          int kick, double* __restrict__ xbar3,
                                                          double* __restrict__ vbar3,
                                                          double* __restrict__ mbar,
                                                          T4* __restrict__ jrec)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const double4 p = pos[i];
    const double ki = kick == UTCNB_KICK_REFERENCE ? p.w : 1.0;
    double al[3], dota = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double ub = vbar3[3 * (size_t)i + k] + xbar3[3 * (size_t)i + k] * dt;
        vbar3[3 * (size_t)i + k] = damping * ub;
        al[k] = damping * ki * dt * ub;
        if (acc3)
            dota += ub * acc3[3 * (size_t)i + k];
    }
    if (acc3)
        mbar[i] += damping * dt * dota;
    jrec[2 * (size_t)i] = T4((E)p.x, (E)p.y, (E)p.z, (E)p.w);
    jrec[2 * (size_t)i + 1] = T4((E)al[0], (E)al[1], (E)al[2], (E)0);
}

template <typename T4, typename E>
__global__ void jrec_pad_kernel(T4* __restrict__ jrec, int n, int n_pad)
{
    const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pad) {
        jrec[2 * (size_t)i] = T4((E)kPadCoord, (E)kPadCoord, (E)kPadCoord, (E)0);
        jrec[2 * (size_t)i + 1] = T4((E)0, (E)0, (E)0, (E)0);
    }
}

hipError_t launch_loss_seed(const double4* pos, const double* tgt3, int n, double* xbar3,
                            double* vbar3, double* mbar, double* part, double* loss,
                            hipStream_t st)
{
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(loss_seed_kernel, dim3(blocks), dim3(256), 0, st, pos, tgt3, n, xbar3,
                       vbar3, mbar, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(loss_finish_kernel, dim3(1), dim3(64), 0, st, part, blocks, n, loss);
    return hipGetLastError();
}

hipError_t launch_adjoint_pre(bool fp32_records, const double4* pos, const double* acc3, int n,
                              double dt, double damping, int kick, double* xbar3, double* vbar3,
                              double* mbar, void* jrec, hipStream_t st)
{
    const int blocks = (n + 255) / 256;
    if (fp32_records)
        hipLaunchKernelGGL((adjoint_pre_kernel<float4, float>), dim3(blocks), dim3(256), 0, st, pos,
                           acc3, n, dt, damping, kick, xbar3, vbar3, mbar, (float4*)jrec);
    else
        hipLaunchKernelGGL((adjoint_pre_kernel<double4, double>), dim3(blocks), dim3(256), 0, st, pos,
                           acc3, n, dt, damping, kick, xbar3, vbar3, mbar, (double4*)jrec);
    return hipGetLastError();
}

hipError_t launch_jrec_pad(bool fp32_records, void* jrec, int n, int n_pad, hipStream_t st)
{
    if (n_pad <= n)
        return hipSuccess;
    const int cnt = n_pad - n;
    if (fp32_records)
        hipLaunchKernelGGL((jrec_pad_kernel<float4, float>), dim3((cnt + 255) / 256), dim3(256), 0, st,
                           (float4*)jrec, n, n_pad);
    else
        hipLaunchKernelGGL((jrec_pad_kernel<double4, double>), dim3((cnt + 255) / 256), dim3(256), 0, st,
                           (double4*)jrec, n, n_pad);
    return hipGetLastError();
}

}  // namespace utcnb
