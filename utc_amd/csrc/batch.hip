// batch.hip -- many independent systems of uneven size in persistent kernels.
//
// The reference runs independent systems as independent UTC tasks, each with
// its own worker thread and stream (benchapps/nbody/utc/nbody_mtask.cc:185-231).
// Here all systems live in one concatenated allocation (each padded to a
// multiple of 64 bodies with far-away zero-mass bodies) and one launch per
// step covers every system:
//   * a work item is (system s, target tile of 64 bodies); W waves of the
//     workgroup split that system's sources into W contiguous runs of 64-body
//     chunks and meet in LDS, like the single-system kernel;
//   * the grid is persistent (as many workgroups as stay resident); the host
//     assigns items to workgroups once, longest first onto the least loaded
//     workgroup (LPT), so every workgroup walks a fixed list;
//   * the step (force + FP64 kick-drift + snapshot) and the adjoint's VJP use
//     the same item lists; the whole forward (+ adjoint) sequence is captured
//     once in a HIP graph and replayed.
#include "force_tmpl.h"
#include "vjp_pair.h"

namespace utcnb {

template <int W>
__global__ __launch_bounds__(64 * W) void batch_step_kernel(BatchGeom g, const float4* __restrict__ snap_old,
                                                            float4* __restrict__ snap_new,
                                                            double4* __restrict__ pos,
                                                            double4* __restrict__ vel, float eps2,
                                                            double dt, double damping, int kick)
{
    __shared__ float red[W][6][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int it0 = g.start[blockIdx.x], it1 = g.start[blockIdx.x + 1];
    for (int it = it0; it < it1; ++it) {
        const int2 w = g.items[it];
        const int off = g.off[w.x], n = g.n[w.x], C = g.chunks[w.x];
        const int t = w.y * 64 + lane;
        const float4 p = snap_old[off + (t < n ? t : n - 1)];
        float px[1] = {p.x}, py[1] = {p.y}, pz[1] = {p.z};
        float hx[1] = {0.f}, hy[1] = {0.f}, hz[1] = {0.f}, cx[1] = {0.f}, cy[1] = {0.f}, cz[1] = {0.f};
        const int c0 = (int)((long long)C * wave / W), c1 = (int)((long long)C * (wave + 1) / W);
        stream_sources<1, true, false>(snap_old + off, c0 * kChunk, c1 * kChunk, eps2, px, py, pz, hx,
                                       hy, hz, cx, cy, cz);
        red[wave][0][lane] = hx[0];
        red[wave][1][lane] = hy[0];
        red[wave][2][lane] = hz[0];
        red[wave][3][lane] = cx[0];
        red[wave][4][lane] = cy[0];
        red[wave][5][lane] = cz[0];
        __syncthreads();
        if (wave == 0) {
            for (int q = 1; q < W; ++q) {
                kahan_add(hx[0], cx[0], red[q][0][lane] - red[q][3][lane]);
                kahan_add(hy[0], cy[0], red[q][1][lane] - red[q][4][lane]);
                kahan_add(hz[0], cz[0], red[q][2][lane] - red[q][5][lane]);
            }
            if (t < n) {
                EpiMixed epi{pos, vel, snap_new, dt, damping, kick};
                epi(off + t, hx[0], hy[0], hz[0], cx[0], cy[0], cz[0]);
            }
        }
        __syncthreads();
    }
}

// One chunk of 64 sources of the VJP (same algebra as vjp_chunk in adjoint.hip).
// EXACT: the chunk may hold j = i (vjp_pair.h).
template <bool EXACT>
__device__ __forceinline__ void batch_vjp_chunk(const float4* __restrict__ src, int jc, const float4& pi,
                                                const float4& ai, float eps2, float p[4])
{
#pragma unroll 2
    for (int j = jc; j < jc + kChunk; ++j) {
        const float4 pj = src[2 * j];
        const float4 aj = src[2 * j + 1];
        float rx = pj.x - pi.x, ry = pj.y - pi.y, rz = pj.z - pi.z;
        float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
        float y = __builtin_amdgcn_rsqf(d2);
        float y2 = y * y;
        float y3 = y2 * y;
        float y5 = y3 * y2;
        float wx = w_term<EXACT>(pi.w, aj.x, pj.w, ai.x);
        float wy = w_term<EXACT>(pi.w, aj.y, pj.w, ai.y);
        float wz = w_term<EXACT>(pi.w, aj.z, pj.w, ai.z);
        float rw = __builtin_fmaf(rx, wx, __builtin_fmaf(ry, wy, rz * wz));
        float q = -3.0f * y5 * rw;
        p[0] = __builtin_fmaf(q, rx, __builtin_fmaf(y3, wx, p[0]));
        p[1] = __builtin_fmaf(q, ry, __builtin_fmaf(y3, wy, p[1]));
        p[2] = __builtin_fmaf(q, rz, __builtin_fmaf(y3, wz, p[2]));
        float ar = __builtin_fmaf(aj.x, rx, __builtin_fmaf(aj.y, ry, aj.z * rz));
        p[3] = __builtin_fmaf(-y3, ar, p[3]);
    }
}

// Adjoint VJP over items (same algebra as vjp_kernel in adjoint.hip).
template <int W>
__global__ __launch_bounds__(64 * W) void batch_vjp_kernel(BatchGeom g, const float4* __restrict__ jrec,
                                                           float eps2, double* __restrict__ xbar3,
                                                           double* __restrict__ mbar)
{
    __shared__ float red[W][8][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int it0 = g.start[blockIdx.x], it1 = g.start[blockIdx.x + 1];
    for (int it = it0; it < it1; ++it) {
        const int2 w = g.items[it];
        const int off = g.off[w.x], n = g.n[w.x], C = g.chunks[w.x];
        const int t = w.y * 64 + lane;
        const int il = off + (t < n ? t : n - 1);
        const float4 pi = jrec[2 * il];
        const float4 ai = jrec[2 * il + 1];
        float gs[4] = {0, 0, 0, 0}, cs[4] = {0, 0, 0, 0};
        const int c0 = (int)((long long)C * wave / W), c1 = (int)((long long)C * (wave + 1) / W);
        const float4* src = jrec + 2 * (size_t)off;
        for (int jc = c0 * kChunk; jc < c1 * kChunk; jc += kChunk) {
            float p[4] = {0, 0, 0, 0};
            if (jc == w.y * 64)  // the chunk holding j = i for every lane
                batch_vjp_chunk<true>(src, jc, pi, ai, eps2, p);
            else
                batch_vjp_chunk<false>(src, jc, pi, ai, eps2, p);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                kahan_add(gs[q], cs[q], p[q]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            red[wave][q][lane] = gs[q];
            red[wave][4 + q][lane] = cs[q];
        }
        __syncthreads();
        if (wave == 0) {
            for (int o = 1; o < W; ++o)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    kahan_add(gs[q], cs[q], red[o][q][lane] - red[o][4 + q][lane]);
            if (t < n) {
                const size_t i = (size_t)(off + t);
                xbar3[3 * i + 0] += (double)gs[0] - (double)cs[0];
                xbar3[3 * i + 1] += (double)gs[1] - (double)cs[1];
                xbar3[3 * i + 2] += (double)gs[2] - (double)cs[2];
                mbar[i] += (double)gs[3] - (double)cs[3];
            }
        }
        __syncthreads();
    }
}

// Per-system loss: one workgroup per system.  Seeds xbar = 2(x - t)/n_s for
// real bodies, zeroes xbar/vbar/mbar of every body of the system (padding
// included) and writes loss[s] = (1/n_s) sum |x - t|^2 (deterministic order).
__global__ __launch_bounds__(256) void batch_loss_kernel(BatchGeom g, const double4* __restrict__ pos,
                                                         const double* __restrict__ tgt3,
                                                         double* __restrict__ xbar3,
                                                         double* __restrict__ vbar3,
                                                         double* __restrict__ mbar,
                                                         double* __restrict__ loss)
{
    const int s = blockIdx.x;
    const int off = g.off[s], n = g.n[s], np = g.chunks[s] * kChunk;
    double acc = 0.0, c = 0.0;
    for (int t = threadIdx.x; t < np; t += 256) {
        const size_t i = (size_t)(off + t);
        double d[3] = {0, 0, 0};
        if (t < n) {
            const double4 p = pos[i];
            d[0] = p.x - tgt3[3 * i + 0];
            d[1] = p.y - tgt3[3 * i + 1];
            d[2] = p.z - tgt3[3 * i + 2];
            double y = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - c;
            double u = acc + y;
            c = (u - acc) - y;
            acc = u;
        }
        const double f = 2.0 / n;
        for (int k = 0; k < 3; ++k) {
            xbar3[3 * i + k] = f * d[k];
            vbar3[3 * i + k] = 0.0;
        }
        mbar[i] = 0.0;
    }
    double v = acc - c;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v += __shfl_xor(v, o, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        loss[s] = (((red[0] + red[1]) + red[2]) + red[3]) / n;
}

// Padding bodies of every system (far away, zero mass), FP64 and FP32 copies.
__global__ void batch_pad_kernel(BatchGeom g, int nsys, double4* pos, double4* vel, float4* s0,
                                 float4* s1)
{
    const int s = blockIdx.x;
    if (s >= nsys)
        return;
    const int off = g.off[s], n = g.n[s], np = g.chunks[s] * kChunk;
    for (int t = n + threadIdx.x; t < np; t += blockDim.x) {
        pos[off + t] = make_double4(kPadCoord, kPadCoord, kPadCoord, 0.0);
        vel[off + t] = make_double4(0.0, 0.0, 0.0, 1.0);
        s0[off + t] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
        s1[off + t] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
    }
}

hipError_t launch_batch_step(int W, const BatchGeom& g, int nblocks, const float4* snap_old,
                             float4* snap_new, double4* pos, double4* vel, float eps2, double dt,
                             double damping, int kick, hipStream_t st)
{
    switch (W) {
    case 4:
        hipLaunchKernelGGL(batch_step_kernel<4>, dim3(nblocks), dim3(256), 0, st, g, snap_old,
                           snap_new, pos, vel, eps2, dt, damping, kick);
        break;
    case 8:
        hipLaunchKernelGGL(batch_step_kernel<8>, dim3(nblocks), dim3(512), 0, st, g, snap_old,
                           snap_new, pos, vel, eps2, dt, damping, kick);
        break;
    case 16:
        hipLaunchKernelGGL(batch_step_kernel<16>, dim3(nblocks), dim3(1024), 0, st, g, snap_old,
                           snap_new, pos, vel, eps2, dt, damping, kick);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_vjp(int W, const BatchGeom& g, int nblocks, const float4* jrec, float eps2,
                            double* xbar3, double* mbar, hipStream_t st)
{
    switch (W) {
    case 4:
        hipLaunchKernelGGL(batch_vjp_kernel<4>, dim3(nblocks), dim3(256), 0, st, g, jrec, eps2,
                           xbar3, mbar);
        break;
    case 8:
        hipLaunchKernelGGL(batch_vjp_kernel<8>, dim3(nblocks), dim3(512), 0, st, g, jrec, eps2,
                           xbar3, mbar);
        break;
    case 16:
        hipLaunchKernelGGL(batch_vjp_kernel<16>, dim3(nblocks), dim3(1024), 0, st, g, jrec, eps2,
                           xbar3, mbar);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_loss(const BatchGeom& g, int nsys, const double4* pos, const double* tgt3,
                             double* xbar3, double* vbar3, double* mbar, double* loss,
                             hipStream_t st)
{
    hipLaunchKernelGGL(batch_loss_kernel, dim3(nsys), dim3(256), 0, st, g, pos, tgt3, xbar3, vbar3,
                       mbar, loss);
    return hipGetLastError();
}

hipError_t launch_batch_pad(const BatchGeom& g, int nsys, double4* pos, double4* vel, float4* s0,
                            float4* s1, hipStream_t st)
{
    hipLaunchKernelGGL(batch_pad_kernel, dim3(nsys), dim3(64), 0, st, g, nsys, pos, vel, s0, s1);
    return hipGetLastError();
}

}  // namespace utcnb
