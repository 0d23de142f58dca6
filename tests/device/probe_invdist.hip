// probe_invdist.hip -- exhaustive device check of ref_inv_dist (ref_inv.h).
//
// For every float d2 from 2^-96 to FLT_MAX, +inf and a quiet NaN, compares
// ref_inv_dist(d2) bit for bit with (float)(1.0 / (double)sqrtf(d2)) as the
// compiler builds it with IEEE sqrt and division (no fast math), and reports
// the number of mismatches, the first few, and the largest errors in ulps of
// the raw v_rcp_f32 and v_sqrt_f32 approximations over the same inputs
// (against the correctly rounded reciprocal and root: the margins the two
// correction steps start from).  One thread per bit pattern in a
// grid-stride loop over a 32-bit range; counters via global atomics.
// The FP64 variant (ref_inv_dist64) is compared with 1.0 / sqrt(d2) on 2^32
// pseudo-random d2 spread evenly over the binades of [2^-600, 2^600) with
// random significands (splitmix64 of the sample index), plus both ends, and
// the division num / sqrt(d2) (ref_div64 after ref_sqrt64) on 2^31 random
// pairs with num over the binades of [2^-400, 2^400).
// Test infrastructure (run by tests/test_gpu_forward.py).  Build: python -m
// utc_amd.build; run: tests/device/_bin/probe_invdist.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "ref_inv.h"

#define CHK(x)                                                                                     \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);             \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

struct Result {
    unsigned long long bad;
    unsigned long long checked;
    unsigned long long bad64;
    unsigned long long checked64;
    unsigned long long first64;
    unsigned long long baddiv;
    unsigned long long checkeddiv;
    unsigned int first[8];
    unsigned int nfirst;
    unsigned int rcp_err_ulps;   // max |v_rcp_f32(s) - 1.0f / s| in ulps
    unsigned int sqrt_err_ulps;  // max |v_sqrt_f32(x) - sqrtf(x)| in ulps
};

__device__ __attribute__((noinline)) float reference_inv(float d2)
{
    // the reference's expression under IEEE rules (this file has no fast math)
    return (float)(1.0 / (double)sqrtf(d2));
}

__global__ void check(uint32_t lo, uint32_t hi, Result* r)
{
    unsigned long long bad = 0, n = 0;
    unsigned int rcp_ulps = 0, sq_ulps = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; u < (uint64_t)hi; u += stride) {
        const float x = __uint_as_float((uint32_t)u);
        const float want = reference_inv(x);
        const float got = utcnb::ref_inv_dist(x);
        ++n;
        const bool same = __float_as_uint(want) == __float_as_uint(got) || (want != want && got != got);
        if (!same) {
            ++bad;
            const unsigned int k = atomicAdd(&r->nfirst, 1u);
            if (k < 8)
                r->first[k] = (uint32_t)u;
        }
        if (x <= 3.0e38f) {
            const float s = sqrtf(x);
            const int dr = __float_as_int(__builtin_amdgcn_rcpf(s)) - __float_as_int(1.0f / s);
            rcp_ulps = max(rcp_ulps, (unsigned int)(dr < 0 ? -dr : dr));
            const int du = __float_as_int(__builtin_amdgcn_sqrtf(x)) - __float_as_int(s);
            sq_ulps = max(sq_ulps, (unsigned int)(du < 0 ? -du : du));
        }
    }
    atomicAdd(&r->bad, bad);
    atomicAdd(&r->checked, n);
    atomicMax(&r->rcp_err_ulps, rcp_ulps);
    atomicMax(&r->sqrt_err_ulps, sq_ulps);
}

__device__ __attribute__((noinline)) double reference_inv64(double d2)
{
    return 1.0 / sqrt(d2);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Sample k: binade -600 + k % 1200, significand bits from splitmix64(k).
__device__ __forceinline__ double sample64(uint64_t k)
{
    const int e = -600 + (int)(k % 1200);
    const uint64_t mant = splitmix64(k) & 0x000FFFFFFFFFFFFFull;
    const uint64_t bits = ((uint64_t)(e + 1023) << 52) | mant;
    return __longlong_as_double((long long)bits);
}

// Division num / sqrt(d2) (the energy kernel's m_j / |r|): num with a random
// significand over the binades of [2^-400, 2^400), d2 as in check64.
__device__ __attribute__((noinline)) double reference_div64(double num, double d2)
{
    return num / sqrt(d2);
}

__global__ void check64div(uint64_t count, Result* r)
{
    unsigned long long bad = 0, n = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        const double x = sample64(k);
        const uint64_t h = splitmix64(k ^ 0x5DEECE66Dull);
        const int e = -400 + (int)(h % 800);
        const double num = __longlong_as_double((long long)(((uint64_t)(e + 1023) << 52) |
                                                            (splitmix64(h) & 0x000FFFFFFFFFFFFFull)));
        ++n;
        const double want = reference_div64(num, x), got = utcnb::ref_div64(num, utcnb::ref_sqrt64(x));
        if (__double_as_longlong(want) != __double_as_longlong(got)) {
            ++bad;
            atomicMax(&r->first64, (unsigned long long)__double_as_longlong(x));
        }
    }
    atomicAdd(&r->baddiv, bad);
    atomicAdd(&r->checkeddiv, n);
}

__global__ void check64(uint64_t count, Result* r)
{
    unsigned long long bad = 0, n = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count + 2; k += stride) {
        const double x = k < count ? sample64(k) : (k == count ? utcnb::kRefInv64Min : utcnb::kRefInv64Max);
        ++n;
        const double want = reference_inv64(x), got = utcnb::ref_inv_dist64(x);
        if (__double_as_longlong(want) != __double_as_longlong(got)) {
            ++bad;
            atomicMax(&r->first64, (unsigned long long)__double_as_longlong(x));
        }
    }
    atomicAdd(&r->bad64, bad);
    atomicAdd(&r->checked64, n);
}

int main()
{
    Result* d = nullptr;
    CHK(hipMalloc(&d, sizeof(Result)));
    CHK(hipMemset(d, 0, sizeof(Result)));
    const uint32_t lo = __builtin_bit_cast(uint32_t, utcnb::kRefInvMin);
    const uint32_t hi = 0x7F800001u;  // through +inf
    hipLaunchKernelGGL(check, dim3(8192), dim3(256), 0, 0, lo, hi, d);
    CHK(hipGetLastError());
    hipLaunchKernelGGL(check, dim3(1), dim3(64), 0, 0, 0x7FC00000u, 0x7FC00001u, d);  // quiet NaN
    CHK(hipGetLastError());
    const uint64_t count64 = 1ull << 32;
    hipLaunchKernelGGL(check64, dim3(8192), dim3(256), 0, 0, count64, d);
    CHK(hipGetLastError());
    const uint64_t countdiv = 1ull << 31;
    hipLaunchKernelGGL(check64div, dim3(8192), dim3(256), 0, 0, countdiv, d);
    CHK(hipGetLastError());
    Result h{};
    CHK(hipMemcpy(&h, d, sizeof(Result), hipMemcpyDeviceToHost));
    CHK(hipFree(d));
    std::printf("{\"checked\": %llu, \"mismatches\": %llu, \"first\": [", h.checked, h.bad);
    for (unsigned int k = 0; k < h.nfirst && k < 8; ++k)
        std::printf("%s\"0x%08x\"", k ? ", " : "", h.first[k]);
    std::printf("], \"rcp_f32_max_ulps\": %u, \"sqrt_f32_max_ulps\": %u, \"f64_checked\": %llu, "
                "\"f64_mismatches\": %llu, \"f64_largest_mismatch_bits\": \"0x%016llx\", "
                "\"f64div_checked\": %llu, \"f64div_mismatches\": %llu}\n",
                h.rcp_err_ulps, h.sqrt_err_ulps, h.checked64, h.bad64, h.first64, h.checkeddiv, h.baddiv);
    return h.bad == 0 && h.checked == (unsigned long long)(hi - lo) + 1 && h.bad64 == 0 &&
                   h.checked64 == count64 + 2 && h.baddiv == 0 && h.checkeddiv == countdiv
               ? 0
               : 1;
}
