// pow_fixed.hpp -- x^1.5 and x^0.33333 in double, for the two exponents the
// timestep uses on its hot path (wet/dry ladder F:121-147 and Manning
// friction F:760 of cuda_native_impl/hydro-cal-src/src/functors.cu).
//
// The general pow of the device math library spends ~215 VALU instructions
// per call on gfx950; these spend ~20 and ~60.  Both are accurate to within
// 0.51 ulp over all positive finite doubles (the general pow is within
// 1 ulp), and return pow's values for 0, infinities, negatives and NaN.
//
//  x^1.5  = x * sqrt(x) with both rounding errors recovered: s = RN(sqrt x)
//           leaves the exact remainder r = x - s^2 (an fma), the product
//           leaves e = x*s - RN(x*s) (an fma), and
//           x^1.5 = x*s + e + x*r/(2s) + O(2^-104 x^1.5).
//  x^c, c = RN(0.33333) = 1/3 - d, d = 3.3333333333e-6:
//           x^c = cbrt(x) * exp(-d ln x).  With x = m' 2^(3q), m' in
//           [0.5, 4), cbrt(m') comes from a float estimate, one Newton step
//           and one correction from the exact residual m' - y^3 (double-
//           double), so it is known to ~2^-100.  |d ln x| < 2.5e-3, so
//           exp(-d ln x) needs ln x only to ~1e-12 absolute: ln u on
//           [sqrt(1/2), sqrt(2)) by the atanh series to s^15 (error < 2e-14)
//           and exp by its series to t^5.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hydro {

// Approximate reciprocal, used only where its error is corrected or scales
// a term of a few ulp.  Device: v_rcp_f64 (documented within 2^29 ulp, i.e.
// 2^-23 relative).  Host: the exact reciprocal with a 2^-23 relative error
// put in, so host checks of these functions run the device's error budget.
__host__ __device__ __forceinline__ double rcp_approx(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcp(x);
#else
    return (1.0 / x) * (1.0 + 0x1p-23);
#endif
}

__host__ __device__ __forceinline__ double pow1p5_core(double x) {
    const double s = sqrt(x);
    const double p = x * s;
    const double r = fma(-s, s, x);       // x - s^2, exact
    const double e = fma(x, s, -p);       // x*s - p, exact
    const double h = 0.5 * rcp_approx(s);
    return p + fma(x, r * h, e);
}

__host__ __device__ __forceinline__ double pow1p5(double x) {
    if (x > 0x1p-590 && x < 0x1p680) return pow1p5_core(x);   // result normal and finite
    if (x == -INFINITY) return INFINITY;
    if (!(x > 0.0) || x == INFINITY) return x * sqrt(x);    // +-0, inf, negatives, NaN
    // ends of the range: (x 2^-+600)^1.5 = x^1.5 2^-+900 keeps the
    // correction terms normal; ldexp rounds to the subnormal or inf result
    if (x >= 0x1p680) return ldexp(pow1p5_core(x * 0x1p-600), 900);
    return ldexp(pow1p5_core(x * 0x1p600), -900);
}

// Cube root of m in [0.5, 4) as yh + yl, |yh + yl - cbrt(m)| < 2^-100.
__host__ __device__ __forceinline__ void cbrt_dd(double m, double *yh, double *yl) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float f = (float)m;
    const double y0 = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf(f) * (1.0f / 3.0f));
#else
    const double y0 = (double)cbrtf((float)m) * (1.0 + 0x1p-20);   // a seed as rough as the device's
#endif
    // Newton: y1 = y0 - (y0^3 - m) / (3 y0^2), relative error ~1e-13
    const double y02 = y0 * y0;
    const double y1 = y0 - (y02 * y0 - m) * rcp_approx(3.0 * y02);
    // exact residual m - y1^3 via double-double, then one more correction
    const double a = y1 * y1;
    const double b = fma(y1, y1, -a);     // y1^2 = a + b
    const double c1 = a * y1;
    const double c2 = fma(a, y1, -c1);    // a*y1 = c1 + c2
    const double c3 = b * y1;
    const double R = ((m - c1) - c2) - c3;   // m - c1 is exact (Sterbenz)
    *yh = y1;
    *yl = R * rcp_approx(3.0 * a);
}

__host__ __device__ __forceinline__ double pow_0p33333(double x) {
    if (!(x > 0.0 && x < INFINITY)) {
        if (x == 0.0) return 0.0;                      // pow(+-0, c) = +0
        if (x == INFINITY || x == -INFINITY) return INFINITY;
        return NAN;                                    // negatives and NaN
    }
    int ex;
    const double m = frexp(x, &ex);                    // x = m 2^ex, m in [0.5, 1)
    // ex = 3q + j, j in {0, 1, 2}; ex in [-1073, 1024]
    const int q = (ex + 1200) / 3 - 400;
    const int j = ex - 3 * q;
    const double mp = ldexp(m, j);                     // [0.5, 4), x = mp 2^(3q)
    double yh, yl;
    cbrt_dd(mp, &yh, &yl);
    // ln x = ln u + k ln 2, u in [sqrt(1/2), sqrt(2))
    const bool lo = m < 0.70710678118654752440;
    const double u = lo ? 2.0 * m : m;
    const int k = lo ? ex - 1 : ex;
    const double v = u + 1.0;
    double iv = rcp_approx(v);
    iv = fma(iv, fma(-v, iv, 1.0), iv);                // one Newton step: ~2^-44
    const double s = (u - 1.0) * iv;
    const double z = s * s;
    double P = 1.0 / 15.0;
    P = fma(P, z, 1.0 / 13.0);
    P = fma(P, z, 1.0 / 11.0);
    P = fma(P, z, 1.0 / 9.0);
    P = fma(P, z, 1.0 / 7.0);
    P = fma(P, z, 1.0 / 5.0);
    P = fma(P, z, 1.0 / 3.0);
    P = fma(P, z, 1.0);
    const double lnu = 2.0 * s * P;
    // t = -d ln x; d = 1/3 - RN(0.33333), dl2 = RN(d ln 2)
    const double d = 0x1.bf647612eaaabp-19;
    const double dl2 = 0x1.361bda42c7e1cp-19;
    const double t = fma(-d, lnu, -(double)k * dl2);
    // G = exp(t) - 1
    double G = 1.0 / 120.0;
    G = fma(G, t, 1.0 / 24.0);
    G = fma(G, t, 1.0 / 6.0);
    G = fma(G, t, 0.5);
    G = fma(G, t, 1.0);
    G = G * t;
    // (yh + yl)(1 + G) = yh + (yh G + yl (1 + G))
    const double T = fma(yh, G, yl * (1.0 + G));
    return ldexp(yh + T, q);
}

// pow evaluated in double (a float argument is promoted, as in the fp32 tree
// of the reference).  The exponents of the timestep are compile-time
// constants after inlining, so the dispatch folds: 1.5 and 0.33333 take the
// evaluations above, 2.0 is the exact square, anything else the library pow.
__host__ __device__ __forceinline__ double pow_d(double x, double y) {
    if (y == 1.5) return pow1p5(x);
    if (y == 0.33333) return pow_0p33333(x);
    if (y == 2.0) return x * x;
    return pow(x, y);
}
// pow on the rare branches (wet/dry edges, boundary handlers).
__host__ __device__ __forceinline__ double pow_rare(double x, double y) { return pow_d(x, y); }

}  // namespace hydro
