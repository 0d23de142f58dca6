// aux.hip -- O(N) helper kernels (padding, snapshots, conversions) and the
// FP64 energy / momentum diagnostics.
#include "internal.h"
#include "ref_inv.h"

namespace utcnb {

__global__ void pad_f32_kernel(float4* p, int n, int n_pad)
{
    int i = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pad)
        p[i] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
}

__global__ void pad_f64_kernel(double4* p, int n, int n_pad)
{
    int i = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pad)
        p[i] = make_double4(kPadCoord, kPadCoord, kPadCoord, 0.0);
}

hipError_t launch_pad_f32(float4* p, int n, int n_pad, hipStream_t st)
{
    if (n_pad <= n)
        return hipSuccess;
    int cnt = n_pad - n;
    hipLaunchKernelGGL(pad_f32_kernel, dim3((cnt + 255) / 256), dim3(256), 0, st, p, n, n_pad);
    return hipGetLastError();
}

hipError_t launch_pad_f64(double4* p, int n, int n_pad, hipStream_t st)
{
    if (n_pad <= n)
        return hipSuccess;
    int cnt = n_pad - n;
    hipLaunchKernelGGL(pad_f64_kernel, dim3((cnt + 255) / 256), dim3(256), 0, st, p, n, n_pad);
    return hipGetLastError();
}

__global__ void snapshot_kernel(const double4* __restrict__ pos, float4* __restrict__ snap, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double4 p = pos[i];
        snap[i] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
    }
}

hipError_t launch_snapshot(const double4* pos, float4* snap, int n, hipStream_t st)
{
    hipLaunchKernelGGL(snapshot_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pos, snap, n);
    return hipGetLastError();
}

__global__ void f32_to_f64_kernel(const float4* __restrict__ s, double4* __restrict__ d, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float4 v = s[i];
        d[i] = make_double4(v.x, v.y, v.z, v.w);
    }
}

hipError_t launch_f32_to_f64(const float4* src, double4* dst, int n, hipStream_t st)
{
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3((n + 255) / 256), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double4* src, float4* dst, int n, hipStream_t st)
{
    return launch_snapshot(src, dst, n, st);
}

// ---- energy / momentum ---------------------------------------------------------
// Pass 1: one thread per body i computes phi_i = sum_{j != i} m_j / sqrt(r^2+eps2)
// (FP64, compensated chunks), kinetic_i and momentum_i; the workgroup reduces
// its 256 bodies with wavefront shuffles and LDS and writes one partial of 5
// values.  Pass 2: one workgroup adds the partials in a fixed order, so the
// result is deterministic.

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(256) void energy_pass1(const double4* __restrict__ pos,
                                                    const double4* __restrict__ vel, int n,
                                                    double eps2, double* __restrict__ part)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int il = i < n ? i : n - 1;
    const double4 p = pos[il];
    double phi = 0.0, c = 0.0;
    for (int jc = 0; jc < n; jc += kChunk) {
        const int je = jc + kChunk < n ? jc + kChunk : n;
        double a = 0.0;
        int j = jc;
        // four sources per trip, one wave-uniform choice between the short
        // forms of the correctly rounded sqrt and division (ref_inv.h; the
        // same bits inside their ranges) and the generic ones
        for (; j + 4 <= je; j += 4) {
            double d2[4], m[4], t[4];
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double4 q = pos[j + k];
                double rx = q.x - p.x, ry = q.y - p.y, rz = q.z - p.z;
                d2[k] = rx * rx + ry * ry + rz * rz + eps2;
                m[k] = q.w;
                ok = ok && d2[k] >= kRefInv64Min && d2[k] <= kRefInv64Max && m[k] >= kRefDiv64NumMin &&
                     m[k] <= kRefDiv64NumMax;
            }
            if (__ballot(!ok) == 0) {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    t[k] = ref_div64(m[k], ref_sqrt64(d2[k]));
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    t[k] = m[k] / __builtin_sqrt(d2[k]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                a += (j + k == il) ? 0.0 : t[k];
        }
        for (; j < je; ++j) {
            const double4 q = pos[j];
            double rx = q.x - p.x, ry = q.y - p.y, rz = q.z - p.z;
            double d2 = rx * rx + ry * ry + rz * rz + eps2;
            double t = q.w / __builtin_sqrt(d2);
            a += (j == il) ? 0.0 : t;
        }
        double y = a - c;
        double s = phi + y;
        c = (s - phi) - y;
        phi = s;
    }
    double v[5] = {0, 0, 0, 0, 0};
    if (i < n) {
        const double4 u = vel[i];
        v[0] = 0.5 * p.w * (u.x * u.x + u.y * u.y + u.z * u.z);
        v[1] = -0.5 * p.w * (phi - c);
        v[2] = p.w * u.x;
        v[3] = p.w * u.y;
        v[4] = p.w * u.z;
    }
    __shared__ double red[4][5];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double w = wave_sum(v[k]);
        if (lane == 0)
            red[wave][k] = w;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        int k = threadIdx.x;
        part[5 * (size_t)blockIdx.x + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
}

__global__ __launch_bounds__(64) void energy_pass2(const double* __restrict__ part, int nparts,
                                                   double* __restrict__ out5)
{
    const int lane = threadIdx.x;
    for (int k = 0; k < 5; ++k) {
        double s = 0.0, c = 0.0;
        for (int b = lane; b < nparts; b += 64) {
            double y = part[5 * (size_t)b + k] - c;
            double t = s + y;
            c = (t - s) - y;
            s = t;
        }
        double w = wave_sum(s - c);
        if (lane == 0)
            out5[k] = w;
    }
}

// `out5` must hold 5 + 5 * ceil(n/256) doubles: the partials live after the result.
hipError_t launch_energy(const double4* pos, const double4* vel, int n, double eps2,
                         double* out5, hipStream_t st)
{
    const int blocks = (n + 255) / 256;
    double* part = out5 + 5;
    hipLaunchKernelGGL(energy_pass1, dim3(blocks), dim3(256), 0, st, pos, vel, n, eps2, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(energy_pass2, dim3(1), dim3(64), 0, st, part, blocks, out5);
    return hipGetLastError();
}

}  // namespace utcnb
