// Device evaluator of the exponential integral E1(x), x > 0.
//
// The reference evaluates E1 inline with its own rational fits
// (PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_real.cpp:19-48,166-205).
// This evaluator uses the library's own fits (e1_coeffs.h, generated by
// gen_e1_coeffs.py from mpmath):
//   x <  1:  E1(x) = -gamma - ln x + x A(x),  A a polynomial in x,
//   x >= 1:  E1(x) = e^{-x} t P(t)/Q(t),      t = 1/x,
// in accuracy tiers: kE1Full (max rel. error < 1e-18 before rounding) and
// kE1Fast (< 1e-12, for tolerances of 1e-10 and above).
#pragma once

#include <hip/hip_runtime.h>

#include "e1_coeffs.h"
#include "exp_coeffs.h"

namespace se2p {

constexpr int kE1Full = 0;
constexpr int kE1Fast = 1;
// The FP32 tier of the real-space sums: far pairs (x >= 1) in FP32 with the
// 'f32' fit below; near pairs, self terms and the sums' accumulation in FP64
// (near E1 by the fast tier).
constexpr int kF32 = 2;
__host__ __device__ constexpr int fp64_tier(int tier) { return tier == kF32 ? kE1Fast : tier; }

template <int N>
__device__ __forceinline__ double horner(const double (&c)[N], double x)
{
    double r = c[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) r = fma(r, x, c[i]);
    return r;
}

// sum_i c[i] x^(N-1-i): the polynomial with the coefficients in reverse order
template <int N>
__device__ __forceinline__ double horner_rev(const double (&c)[N], double x)
{
    double r = c[0];
#pragma unroll
    for (int i = 1; i < N; ++i) r = fma(r, x, c[i]);
    return r;
}

template <int TIER>
struct E1Coeffs;

template <>
struct E1Coeffs<kE1Full> {
    static constexpr double near[SE2P_E1_FULL_NEAR_DEG + 1] = SE2P_E1_FULL_NEAR;
    static constexpr double farp[SE2P_E1_FULL_FAR_DEG + 1] = SE2P_E1_FULL_FARP;
    static constexpr double farq[SE2P_E1_FULL_FAR_DEG + 1] = SE2P_E1_FULL_FARQ;
};

template <>
struct E1Coeffs<kE1Fast> {
    static constexpr double near[SE2P_E1_FAST_NEAR_DEG + 1] = SE2P_E1_FAST_NEAR;
    static constexpr double farp[SE2P_E1_FAST_FAR_DEG + 1] = SE2P_E1_FAST_FARP;
    static constexpr double farq[SE2P_E1_FAST_FAR_DEG + 1] = SE2P_E1_FAST_FARQ;
};

// FP32 far fit E1(x) = e^{-x} P~(x)/(x Q~(x)) with the reversed polynomials
// (relative error 2e-8 before rounding).
struct E1F32 {
    static constexpr float farp[SE2P_E1_F32_FAR_DEG + 1] = SE2P_E1_F32_FARP;
    static constexpr float farq[SE2P_E1_F32_FAR_DEG + 1] = SE2P_E1_F32_FARQ;
};

template <int N>
__device__ __forceinline__ float horner_rev_f(const float (&c)[N], float x)
{
    float r = c[0];
#pragma unroll
    for (int i = 1; i < N; ++i) r = fmaf(r, x, c[i]);
    return r;
}

// e^{-x} for 0 <= x < 700: Cody-Waite reduction r = n ln 2 - x with n =
// rint(x log2 e), |r| <= ln(2)/2, e^r by the degree-10 Chebyshev-interpolated
// polynomial of exp_coeffs.h (max rel. error 3e-16 before rounding), scaled by
// 2^{-n}.  No special-case handling: the callers pass finite x in range.
// a * b + c as one three-operand v_fma_f64.  The compiler otherwise turns each
// Horner step on coefficients held in vector registers into a copy plus a
// two-operand v_fmac_f64, twice the instructions.
__device__ __forceinline__ double fma3(double a, double b, double c)
{
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

__device__ __forceinline__ double exp_neg(double x)
{
    constexpr double c[SE2P_EXP_DEG + 1] = SE2P_EXP_POLY;
    const double n = __builtin_rint(x * SE2P_LOG2E);
    const double r = fma(n, SE2P_LN2_LO, fma(n, SE2P_LN2_HI, -x));
    double p = c[SE2P_EXP_DEG];
#pragma unroll
    for (int i = SE2P_EXP_DEG - 1; i >= 0; --i) p = fma3(p, r, c[i]);
    return ldexp(p, -(int)n);
}

// 1/x from the hardware reciprocal estimate and two Newton steps (about 1 ulp).
__device__ __forceinline__ double rcp64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double r = fma(-x, y, 1.0);
    y = fma(y, r, y);
    r = fma(-x, y, 1.0);
    return fma(y, r, y);
}

// 1/x to about 1e-14 relative: the hardware estimate (about 2^-23) and one
// Newton step; for the fast tier, whose fits carry 1e-12.
__device__ __forceinline__ double rcp64_1(double x)
{
    const double y = __builtin_amdgcn_rcp(x);
    return fma(y, fma(-x, y, 1.0), y);
}

// The two regimes separately: near for x < 1, far (given e = exp(-x), t = 1/x) for x >= 1.
// ln(x) for finite x > 0 (normal or not): x = m 2^k with m in [sqrt(1/2),
// sqrt(2)), ln(m) = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716, summed as
// 2 s sum_j s^(2j) / (2j + 1) for j <= 10 (the series' tail is below 3e-18
// relative), then k ln 2 with ln 2 split into a 32-bit head and a tail (the
// reduction of W. J. Cody and W. Waite, Software Manual for the Elementary
// Functions, 1980).  About 25 instructions; within 2 ulp of ln.
__device__ __forceinline__ double log_pos(double x)
{
    int k = __builtin_amdgcn_frexp_exp(x);
    double m = __builtin_amdgcn_frexp_mant(x);  // [1/2, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        k -= 1;
    }
    const double s = (m - 1.0) * rcp64(m + 1.0);
    const double z = s * s;
    double p = 1.0 / 21;
    p = fma(p, z, 1.0 / 19);
    p = fma(p, z, 1.0 / 17);
    p = fma(p, z, 1.0 / 15);
    p = fma(p, z, 1.0 / 13);
    p = fma(p, z, 1.0 / 11);
    p = fma(p, z, 1.0 / 9);
    p = fma(p, z, 1.0 / 7);
    p = fma(p, z, 1.0 / 5);
    p = fma(p, z, 1.0 / 3);
    const double lm = 2.0 * fma(s * z, p, s);
    const double kd = (double)k;
    return fma(kd, SE2P_LN2_HI, fma(kd, SE2P_LN2_LO, lm));
}

// ln(x) for normal x > 0 to about 1e-15 absolute: x = m 2^k with m in [1/2, 1),
// y0 = ln m from the single-precision hardware log2 (absolute error about
// 1e-7), one Newton step for e^y = m, y1 = y0 + m e^{-y0} - 1 (quadratic: the
// error becomes about y0's squared over 2), then k ln 2 with the split ln 2.
// The exponential is exp_neg, whose coefficients the caller already holds in
// registers, so this log adds no constants of its own (log_pos adds ten).
__device__ __forceinline__ double log_newton(double x)
{
    const int k = __builtin_amdgcn_frexp_exp(x);
    const double m = __builtin_amdgcn_frexp_mant(x);
    const double y0 = (double)__builtin_amdgcn_logf((float)m) * SE2P_LN2_HI;
    const double lm = y0 + fma(m, exp_neg(y0), -1.0);
    const double kd = (double)k;
    return fma(kd, SE2P_LN2_HI, fma(kd, SE2P_LN2_LO, lm));
}

template <int TIER>
__device__ __forceinline__ double e1_near(double x)
{
    using C = E1Coeffs<TIER>;
    // the fast tier takes the library's own log (~25 instructions against ~80
    // for the math library's compensated one)
    const double lx = TIER == kE1Fast ? log_newton(x) : log(x);
    return (-SE2P_EULER_GAMMA - lx) + x * horner(C::near, x);
}

template <int TIER>
__device__ __forceinline__ double e1_far(double e, double t)
{
    using C = E1Coeffs<TIER>;
    return e * t * horner(C::farp, t) * rcp64(horner(C::farq, t));
}

// E1(x) given e = exp(-x) and t = 1/x, both already needed by the caller.
template <int TIER>
__device__ __forceinline__ double e1_eval(double x, double e, double t)
{
    if (x < 1.0) return e1_near<TIER>(x);
    return e1_far<TIER>(e, t);
}

}  // namespace se2p
