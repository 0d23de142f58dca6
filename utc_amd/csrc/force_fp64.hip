// force_fp64.hip -- FP64 all-pairs kernels: the on-device yardstick that the
// FP32 and mixed paths are measured against at sizes a CPU cannot reach.
//
// Same mapping as the FP32 kernels (sources uniform per wave, scalar loads of
// double4 bodies into SGPR pairs, S waves splitting the sources), with a
// correctly rounded sqrt and division per pair (ref_inv_dist64 where it gives
// the same bits) and Kahan-compensated chunk sums, so every field component
// carries only a few FP64 ulps of error.
#include "internal.h"
#include "ref_inv.h"

namespace utcnb {

__device__ __forceinline__ void kahan_add64(double& s, double& c, double p)
{
    double y = p - c;
    double t = s + y;
    c = (t - s) - y;
    s = t;
}

struct Epi64Step {
    const double4* pos_old;
    double4* pos_new;
    double4* vel;
    double dt, damping;
    int kick;
    __device__ __forceinline__ void operator()(int i, double ax, double ay, double az) const
    {
        double4 p = pos_old[i];
        double4 v = vel[i];
        double w = kick == UTCNB_KICK_REFERENCE ? p.w : 1.0;
        v.x = (v.x + (ax * w) * dt) * damping;
        v.y = (v.y + (ay * w) * dt) * damping;
        v.z = (v.z + (az * w) * dt) * damping;
        p.x += v.x * dt;
        p.y += v.y * dt;
        p.z += v.z * dt;
        pos_new[i] = p;
        vel[i] = v;
    }
};

// Field only; body i goes to row i - off (off = the first target of a slice).
struct Epi64Accel {
    double* acc3;
    int off = 0;
    __device__ __forceinline__ void operator()(int i, double ax, double ay, double az) const
    {
        const size_t r = 3 * (size_t)(i - off);
        acc3[r + 0] = ax;
        acc3[r + 1] = ay;
        acc3[r + 2] = az;
    }
};

// Targets [i0, i0 + nt) against all n_pad sources.
template <int S, int W, class Epi>
__global__ __launch_bounds__(64 * W) void force64_kernel(const double4* __restrict__ src, int i0, int nt,
                                                         int n_pad, double eps2, Epi epi)
{
    constexpr int T = W / S;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tile = blockIdx.x * T + wave / S;
    const int part = wave % S;
    const int t = tile * 64 + lane;
    const double4 pi = src[i0 + (t < nt ? t : nt - 1)];
    double hx = 0, hy = 0, hz = 0, cx = 0, cy = 0, cz = 0;
    const int len = n_pad / S;
    const int j0 = part * len;
    for (int jc = j0; jc < j0 + len; jc += kChunk) {
        double ax = 0, ay = 0, az = 0;
        for (int j = jc; j < jc + kChunk; j += 4) {
            // four sources: their geometry, one wave-uniform choice of the
            // inverse-distance sequence (ref_inv.h: the generic correctly
            // rounded one's operations when every d2 is in [2^-600, 2^600],
            // so the result bits do not depend on the choice), then the sums
            double r[4][3], d2[4], inv[4], m[4];
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double4 b = src[j + k];
                r[k][0] = b.x - pi.x; r[k][1] = b.y - pi.y; r[k][2] = b.z - pi.z;
                m[k] = b.w;
                d2[k] = __builtin_fma(r[k][0], r[k][0], __builtin_fma(r[k][1], r[k][1], __builtin_fma(r[k][2], r[k][2], eps2)));
                ok = ok && d2[k] >= kRefInv64Min && d2[k] <= kRefInv64Max;
            }
            if (__ballot(!ok) == 0) {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    inv[k] = ref_inv_dist64(d2[k]);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    inv[k] = 1.0 / __builtin_sqrt(d2[k]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double s = m[k] * (inv[k] * inv[k] * inv[k]);
                ax = __builtin_fma(r[k][0], s, ax);
                ay = __builtin_fma(r[k][1], s, ay);
                az = __builtin_fma(r[k][2], s, az);
            }
        }
        kahan_add64(hx, cx, ax);
        kahan_add64(hy, cy, ay);
        kahan_add64(hz, cz, az);
    }
    if constexpr (S > 1) {
        __shared__ double red[W][6][64];
        red[wave][0][lane] = hx;
        red[wave][1][lane] = hy;
        red[wave][2][lane] = hz;
        red[wave][3][lane] = cx;
        red[wave][4][lane] = cy;
        red[wave][5][lane] = cz;
        __syncthreads();
        if (part != 0)
            return;
        for (int q = 1; q < S; ++q) {
            kahan_add64(hx, cx, red[wave + q][0][lane] - red[wave + q][3][lane]);
            kahan_add64(hy, cy, red[wave + q][1][lane] - red[wave + q][4][lane]);
            kahan_add64(hz, cz, red[wave + q][2][lane] - red[wave + q][5][lane]);
        }
    }
    if (t < nt)
        epi(i0 + t, hx - cx, hy - cy, hz - cz);
}

template <int S, class Epi>
static hipError_t launch64_tile(const double4* src, int i0, int nt, int n_pad, double eps2, const Epi& epi,
                                hipStream_t st)
{
    constexpr int W = S < 4 ? 4 : S;
    constexpr int T = W / S;
    const int tiles = (nt + 63) / 64;
    const int blocks = (tiles + T - 1) / T;
    hipLaunchKernelGGL((force64_kernel<S, W, Epi>), dim3(blocks), dim3(64 * W), 0, st, src, i0, nt,
                       n_pad, eps2, epi);
    return hipGetLastError();
}

// Source splits by the system's size n (enough waves to fill the chip for all
// n targets); a slice keeps the split of the whole system, so its sums (and
// results) are bitwise those of the same bodies in a full evaluation.
template <class Epi>
static hipError_t launch64(const double4* src, int i0, int nt, int n, int n_pad, double eps2, const Epi& epi,
                           hipStream_t st)
{
    if (nt < 1)
        return hipErrorInvalidValue;
    const int waves = (n + 63) / 64;
    if (waves >= 4096)
        return launch64_tile<4>(src, i0, nt, n_pad, eps2, epi, st);
    if (waves * 8 >= 4096 * 4)
        return launch64_tile<8>(src, i0, nt, n_pad, eps2, epi, st);
    return launch64_tile<16>(src, i0, nt, n_pad, eps2, epi, st);
}

hipError_t launch_fp64_step(const StepArgs& a, const double4* pos_old, double4* pos_new,
                            double4* vel, hipStream_t st)
{
    Epi64Step epi{pos_old, pos_new, vel, a.dt, a.damping, a.kick};
    return launch64(pos_old, 0, a.n, a.n, a.n_pad, a.eps2, epi, st);
}

hipError_t launch_fp64_accel(const StepArgs& a, const double4* src, double* acc3, hipStream_t st)
{
    Epi64Accel epi{acc3};
    return launch64(src, 0, a.n, a.n, a.n_pad, a.eps2, epi, st);
}

hipError_t launch_fp64_step_slice(const StepArgs& a, int i0, int nt, const double4* pos_old, double4* pos_new,
                                  double4* vel, hipStream_t st)
{
    Epi64Step epi{pos_old, pos_new, vel, a.dt, a.damping, a.kick};
    return launch64(pos_old, i0, nt, a.n, a.n_pad, a.eps2, epi, st);
}

hipError_t launch_fp64_accel_slice(const StepArgs& a, int i0, int nt, const double4* src, double* acc3,
                                   hipStream_t st)
{
    Epi64Accel epi{acc3, i0};
    return launch64(src, i0, nt, a.n, a.n_pad, a.eps2, epi, st);
}

}  // namespace utcnb
