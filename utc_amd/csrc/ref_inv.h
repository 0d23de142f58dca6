// ref_inv.h -- the reference's inverse distance (float)(1.0 / (double)sqrtf(d2))
// (benchapps/nbody/seq/bodysystem.h:86) in fewer instructions, bit for bit,
// for d2 >= 2^-96 (device code only).
//
// sqrtf.  v_sqrt_f32 returns s0 within 1 ulp of sqrt(x), so the correctly
// rounded root is s0 or a neighbour.  Tuckerman's test picks it: sqrt(x) <
// s0 - u/2 (u = ulp of s0) iff x < (s0 - u/2)^2 = s0 (s0 - u) + u^2/4.  The
// square of a float midpoint is never a float and x lies on a grid far
// coarser than u^2/4, so x <= s0 (s0 - u) decides, and one FMA gives the sign
// of x - s0 (s0 - u) exactly; likewise x > s0 (s0 + u) for the upper
// neighbour.  (s0 - u and s0 + u are the neighbouring floats: the bit pattern
// minus and plus one.)
//
// 1.0 / (double)s, rounded to float.  1/s stays at least 2^-48 (relative)
// away from every float midpoint (|1 - s m| >= 2^-48 for a float s and a
// midpoint m: s m is a 24-by-25-bit product that cannot equal 1), so the
// FP64 quotient (within 2^-53) rounds to the correctly rounded FP32
// reciprocal of s.  That is computed in FP32: y0 = v_rcp_f32(s) (within 1
// ulp), then one correction y1 = fma(y0, e, y0), e = fma(-s, y0, 1).  From a
// correctly rounded y0 the step returns it for every s; from a y0 one ulp
// off it returns the correctly rounded value except for 33 significands
// (tests/test_ref_inv_cpu.py, in exact arithmetic), so exactness rests on
// v_rcp_f32 not landing there: tests/device/probe_invdist.hip checks every
// float d2 in [2^-96, inf] (every s this path meets) on the device, with no
// mismatch on gfx950.  Against v_rcp_f64, a Newton step in FP64 and two
// conversions, this saves a VALU slot and the FP64 transcendental per pair
// (9% faster reference steps on MI355X).  The clamp of s in the
// residual keeps 1/inf = 0 (an overflowed d2) instead of inf * 0; NaN stays
// NaN.
//
// The compiler's generic sequences rescale tiny inputs; from 2^-96 up they
// are not needed (s and 1/s are normal).  tests/device/probe_invdist.hip
// checks this function against the compiler's correctly rounded sqrtf and
// division for every float d2 in [2^-96, inf] and NaN on the device.
#pragma once

#include <hip/hip_runtime.h>

namespace utcnb {

// Smallest d2 for which ref_inv_dist is exact.
constexpr float kRefInvMin = 0x1p-96f;

__device__ __forceinline__ float ref_sqrt_rn(float x)
{
    const float s0 = __builtin_amdgcn_sqrtf(x);
    const float lo = __int_as_float(__float_as_int(s0) - 1);
    const float hi = __int_as_float(__float_as_int(s0) + 1);
    float s = s0;
    if (__builtin_fmaf(-lo, s0, x) <= 0.f)
        s = lo;
    if (__builtin_fmaf(-hi, s0, x) > 0.f)
        s = hi;
    return s;
}

__device__ __forceinline__ float ref_recip_rn(float s)
{
    const float y0 = __builtin_amdgcn_rcpf(s);
    const float e = __builtin_fmaf(-__builtin_fminf(s, 0x1p127f), y0, 1.0f);
    return __builtin_fmaf(y0, e, y0);
}

// (float)(1.0 / (double)sqrtf(d2)) for d2 >= kRefInvMin (or +inf, NaN).
__device__ __forceinline__ float ref_inv_dist(float d2)
{
    return ref_recip_rn(ref_sqrt_rn(d2));
}

// FP64: 1.0 / sqrt(d2) (benchapps/nbody/seq/bodysystem.h:86 for T = double)
// for d2 in [kRefInv64Min, kRefInv64Max].  The generic correctly rounded
// sequences of this target are, for such arguments, the operations below plus
// steps that then change nothing: sqrt = Goldschmidt's iteration from
// v_rsq_f64 (g = x y, h = y/2, r = 1/2 - h g, g += g r, h += h r) and two
// Newton corrections g += (x - g^2) h, with a rescaling of x below 2^-767 and
// a special-value select for 0 and inf; the division 1/g = v_rcp_f64 and two
// Newton steps z += z (1 - g z), q = z and the residual correction
// q + (1 - g q) z, with v_div_scale / v_div_fmas / v_div_fixup, which rescale
// only for extreme exponent gaps and handle special values.  Between 2^-600
// and 2^600, g and 1/g stay hundreds of binades from every such threshold, so
// the same operations in the same order give the same bits.
// tests/device/probe_invdist.hip compares the two on the device.
constexpr double kRefInv64Min = 0x1p-600, kRefInv64Max = 0x1p600;

// sqrt(x) for x in [kRefInv64Min, kRefInv64Max]: the generic sequence's
// operations without its rescaling and special-value select.
__device__ __forceinline__ double ref_sqrt64(double x)
{
#pragma clang fp contract(off)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return g;
}

// num / den for den = ref_sqrt64(x) of such an x and num in
// [kRefDiv64NumMin, kRefDiv64NumMax] (exponent gap under 700 binades, quotient
// normal): the generic division's operations without v_div_scale /
// v_div_fmas / v_div_fixup, which change nothing there.
constexpr double kRefDiv64NumMin = 0x1p-400, kRefDiv64NumMax = 0x1p400;

__device__ __forceinline__ double ref_div64(double num, double den)
{
#pragma clang fp contract(off)
    double z = __builtin_amdgcn_rcp(den);
    double e = __builtin_fma(-den, z, 1.0);
    z = __builtin_fma(z, e, z);
    e = __builtin_fma(-den, z, 1.0);
    z = __builtin_fma(z, e, z);
    const double q = num * z;
    const double res = __builtin_fma(-den, q, num);
    return __builtin_fma(res, z, q);
}

__device__ __forceinline__ double ref_inv_dist64(double x)
{
    return ref_div64(1.0, ref_sqrt64(x));
}

}  // namespace utcnb
