// probe_valu.hip -- FP32 VALU issue rate of gfx950 without packed FP32.
//
// Every lane runs 8 independent FMA chains (64 FMAs per loop trip), alone or
// with a v_rsq_f32 or a DPP wave rotation per 8 FMAs, 16 waves per SIMD.  The
// rate of the plain chains is the non-packed FP32 issue ceiling the pair
// kernels are measured against (256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz =
// 39.3e12 lane-FMAs/s).  Each lane writes one float; the grid matches the
// output buffer.  Build: python -m utc_amd.build; run: utc_amd/lib/probe_valu.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHK(x)                                                                                     \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);             \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

__device__ __forceinline__ float ror1(float v)
{
    const int x = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x13C /* wave_ror:1 */, 0xF, 0xF, false));
}

// MODE 0: FMA chains only; 1: + one rsq of a chain value per 8 FMAs, summed
// into a separate accumulator; 2: + one DPP rotation of a chain value per 8.
template <int MODE>
__global__ __launch_bounds__(256) void probe(float* out, int iters, float a, float b)
{
    float c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        c[k] = threadIdx.x * 1e-3f + k + 1.0f;
    float acc = 0.f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                c[k] = __builtin_fmaf(c[k], a, b);
            if (MODE == 1)
                acc += __builtin_amdgcn_rsqf(c[u]);
            if (MODE == 2)
                acc += ror1(c[u]);
        }
    }
    float s = acc;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        s += c[k];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <int MODE>
int run(const char* name, int blocks, int iters, float* d)
{
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(probe<MODE>, dim3(blocks), dim3(256), 0, 0, d, iters, 0.999f, 1e-3f);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0));
    hipLaunchKernelGGL(probe<MODE>, dim3(blocks), dim3(256), 0, 0, d, iters, 0.999f, 1e-3f);
    CHK(hipEventRecord(e1));
    CHK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double fma = (double)blocks * 256 * iters * 64;
    std::printf("%-22s %8.3f ms  %.4e lane-FMA/s  %5.1f%% of 39.32e12\n", name, ms, fma / (ms * 1e-3),
                100.0 * fma / (ms * 1e-3) / 39.32e12);
    CHK(hipEventDestroy(e0));
    CHK(hipEventDestroy(e1));
    return 0;
}

int main()
{
    int cus = 0;
    CHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const int grid = cus * 16;  // 16 workgroups of 4 waves per CU: 16 waves per SIMD over the run
    float* d = nullptr;
    CHK(hipMalloc(&d, sizeof(float) * grid * 256));
    const int iters = 20000;
    std::printf("CUs %d, %d workgroups of 256, %d loop trips of 64 FMAs per lane\n", cus, grid, iters);
    if (run<0>("fma", grid, iters, d) || run<1>("fma + rsq per 8", grid, iters, d) ||
        run<2>("fma + dpp per 8", grid, iters, d) || run<0>("fma (again)", grid, iters, d))
        return 1;
    CHK(hipFree(d));
    return 0;
}
