// FP64 elementary functions of the fractional kernels (agwn_frac.hip, agwn_robust.hip), written for the cases they
// meet: finite arguments, results good to ~2 ulp, no errno or exception semantics.
#pragma once

#include <hip/hip_runtime.h>

namespace agwn
{
// atan2 in FP64 for finite arguments, C semantics (signed zeros included).
// Reduction: both arguments are scaled by one power of two so that max(|x|, |y|) lies in [1/2, 1); r = min / max lies
// in [0, 1] and is reduced to |t| <= tan(pi/16) + 1e-15 by one division whose numerator and denominator are picked by
// comparisons: r <= tan(pi/16): t = min / max; r <= tan(3pi/16): t = (min - k max) / (max + k min), k = tan(pi/8)
// rounded to double, atan r = atan k + atan t; else t = (min - max) / (min + max), atan r = pi/4 + atan t (the
// subtraction is exact: min >= max / 2 there). atan t = t P(t^2), P the degree-7 polynomial (odd degree 15 in t)
// fitted by the Remez exchange to the relative error of atan on |t| <= 0.2 (csrc/atan_fit.py: levelled error 1.0e-17,
// 1.62 ulp after rounding to double and Horner evaluation). The division is a reciprocal refined by two Newton steps
// and a residual correction. Quadrant fix-ups: pi/2 - a when |y| > |x|, pi - a when x is negative (or -0), then the
// sign of y.
__device__ __forceinline__ double atan2_f64(double y, double x)
{
    constexpr double c0 = 1.0, c1 = -0.3333333333333009, c2 = 0.19999999998291643, c3 = -0.14285713942546116,
                     c4 = 0.11111077206208915, c5 = -0.09089084924181011, c6 = 0.0763762781320867,
                     c7 = -0.05801001224879643;
    constexpr double k16 = 0.19891236737965800691;  // tan(pi/16)
    constexpr double k316 = 0.66817863791929891999; // tan(3 pi/16)
    constexpr double k8 = 0.41421356237309504880;   // tan(pi/8) (rounded; the offset below is atan of this double)
    constexpr double atan_k8 = 0.39269908169872414;  // atan(double(tan(pi/8))), rounded
    constexpr double pi4 = 0.78539816339744830962, pi2 = 1.5707963267948966192, pi = 3.1415926535897932385;
    double const ax = __builtin_fabs(x), ay = __builtin_fabs(y);
    double mx = __builtin_fmax(ax, ay), mn = __builtin_fmin(ax, ay);
    int const e = __builtin_amdgcn_frexp_exp(mx); // 0 for mx = 0
    mx = __builtin_amdgcn_ldexp(mx, -e);
    mn = __builtin_amdgcn_ldexp(mn, -e);
    bool const lo = mn <= k16 * mx, mid = !lo && mn <= k316 * mx;
    // num = mn, mn - k8 mx or mn - mx and den = mx, mx + k8 mn or mn + mx as one FMA each with k = 0, k8 or 1 (the
    // same single roundings; selects of constants, no divergent branches)
    double const k = lo ? 0.0 : (mid ? k8 : 1.0);
    double const num = __builtin_fma(-k, mx, mn);
    double const den = __builtin_fma(k, mn, mx); // in [1/2, 2) unless both are zero
    double r = __builtin_amdgcn_rcp(den);
    r = __builtin_fma(__builtin_fma(-den, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-den, r, 1.0), r, r);
    double t = num * r;
    t = __builtin_fma(__builtin_fma(-den, t, num), r, t);
    t = mx > 0.0 ? t : 0.0; // atan2(+-0, +-0)
    double const u = t * t;
    double p = c7;
    p = __builtin_fma(p, u, c6);
    p = __builtin_fma(p, u, c5);
    p = __builtin_fma(p, u, c4);
    p = __builtin_fma(p, u, c3);
    p = __builtin_fma(p, u, c2);
    p = __builtin_fma(p, u, c1);
    p = __builtin_fma(p, u, c0);
    double a = __builtin_fma(t, p, lo ? 0.0 : (mid ? atan_k8 : pi4));
    a = ay > ax ? pi2 - a : a;
    a = __builtin_signbit(x) ? pi - a : a;
    return __builtin_copysign(a, y);
}

// sqrt in FP64 of a finite q >= 0 (a squared length), within ~1 ulp: the hardware reciprocal square root refined by
// two coupled Newton-Raphson (Goldschmidt) steps for g = sqrt(q) and h = 1 / (2 sqrt(q)); q below 2^-1000 is scaled
// by 2^256 first (exact), and 0 gives 0.
__device__ __forceinline__ double sqrt_f64(double q)
{
    bool const tiny = q < 0x1p-1000;
    double const qs = tiny ? q * 0x1p256 : q;
    double const r = __builtin_amdgcn_rsq(qs);
    double g = qs * r, h = 0.5 * r;
    double e = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    e = __builtin_fma(-g, g, qs); // residual
    g = __builtin_fma(e, h, g);
    g = tiny ? g * 0x1p-128 : g;
    return qs > 0.0 ? g : 0.0;
}

// sqrt_f64 for q = 0 or q >= 2^-1000 (a squared length of integer coordinates, say): the same steps without the
// scaling of tiny arguments.
__device__ __forceinline__ double sqrt_f64_normal(double q)
{
    double const r = __builtin_amdgcn_rsq(q);
    double g = q * r, h = 0.5 * r;
    double e = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    e = __builtin_fma(-g, g, q); // residual
    g = __builtin_fma(e, h, g);
    return q > 0.0 ? g : 0.0;
}

// a / b in FP64 for finite a and b whose magnitudes and quotient stay well inside the normal range (no scaling steps),
// within ~1 ulp: the hardware reciprocal refined by two Newton steps, then one residual correction of the quotient.
__device__ __forceinline__ double div_f64_normal(double a, double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    double const q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
}

// a + b = s + e exactly (Knuth's two-sum)
__device__ __forceinline__ void two_sum_f64(double a, double b, double& s, double& e)
{
    s = a + b;
    double const bv = s - a;
    double const av = s - bv;
    e = (a - av) + (b - bv);
}
} // namespace agwn
