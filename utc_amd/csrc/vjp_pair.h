// vjp_pair.h -- the w term of the adjoint pair sum (device code only).
//
// The VJP kernels sum over every source, j = i included, as the force kernels
// do.  Their fast form of w = m_i a_j - m_j a_i is fma(-m_j, a_i, m_i * a_j):
// at j = i that is the rounding residual of m_i a_i, not 0, and with r = 0 the
// pair adds J(0) w = eps^-3 w, up to ~6e-8 |m_i a_i| / eps^3 in FP32: more
// than the whole pair sum when eps is small.  The source chunks that can hold
// j = i (the diagonal chunk of a wave) therefore use w_exact: both products
// rounded, no contraction, so w is exactly 0 at j = i.  It costs one more
// VALU op per component and runs on a 1/(N/64) share of the pairs.
#pragma once

#include <hip/hip_runtime.h>

namespace utcnb {

__device__ __forceinline__ float w_fast(float mi, float aj, float mj, float ai)
{
    return __builtin_fmaf(-mj, ai, mi * aj);
}
__device__ __forceinline__ double w_fast(double mi, double aj, double mj, double ai)
{
    return __builtin_fma(-mj, ai, mi * aj);
}

__device__ __forceinline__ float w_exact(float mi, float aj, float mj, float ai)
{
#pragma clang fp contract(off)
    return mi * aj - mj * ai;
}
__device__ __forceinline__ double w_exact(double mi, double aj, double mj, double ai)
{
#pragma clang fp contract(off)
    return mi * aj - mj * ai;
}

template <bool EXACT, typename T>
__device__ __forceinline__ T w_term(T mi, T aj, T mj, T ai)
{
    if constexpr (EXACT)
        return w_exact(mi, aj, mj, ai);
    else
        return w_fast(mi, aj, mj, ai);
}

}  // namespace utcnb
