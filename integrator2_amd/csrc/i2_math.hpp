// i2_math.hpp -- FP64 logarithm and two-argument arctangent for the
// quadrature hot loop.
//
// The library versions cost about 92 (log) and 99 (atan2) VALU instructions
// on gfx950; the inner integral needs three logarithms and one atan2 per
// quadrature point, so they dominate the kernel.  These versions use one
// division each and a short polynomial, are branch-free on the normal path,
// and stay within about 2 ulp (checked on the host against glibc and
// high-precision references).
//
// Methods (textbook):
//   log:   x = 2^k m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s) with
//          s = (m - 1)/(m + 1), |s| <= 0.1716; atanh(s) = s (1 + s^2/3 + s^4/5 + ...),
//          truncated after s^18/19 (truncation below 2^-54 relative);
//          log x = k ln2_hi + (2 atanh s + k ln2_lo), ln2 split so that k ln2_hi is exact.
//   atan2: reduce to t = min(|x|,|y|)/max(|x|,|y|) in [0, 1]; with a centre
//          c in {0, c8, 1} (c8 = the double nearest tan(pi/8)),
//          atan t = atan c + atan u, u = (n - c d)/(d + c n) for t = n/d, so
//          |u| <= tan(pi/16); atan u by its Taylor series to u^21/21;
//          then the octant and quadrant identities.
// Non-finite, zero, negative and subnormal arguments take the library path.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#ifndef I2_HD
#define I2_HD __host__ __device__
#endif

namespace i2m {

constexpr double LN2_HI = 0.69314718036912382;         // ln 2 to 32 significant bits
constexpr double LN2_LO = 1.9082149292705877e-10;      // ln 2 - LN2_HI
constexpr double SQRT_HALF = 0.70710678118654757;
constexpr double PI_HI = 3.1415926535897931, PI_LO = 1.2246467991473532e-16;
constexpr double PIO2_HI = 1.5707963267948966, PIO2_LO = 6.123233995736766e-17;
constexpr double PIO4_HI = 0.78539816339744828, PIO4_LO = 3.061616997868383e-17;
constexpr double C8 = 0.41421356237309503;             // double nearest tan(pi/8)
constexpr double ATAN_C8_HI = 0.39269908169872414;     // atan(C8), split
constexpr double ATAN_C8_LO = 3.060132146563891e-18;
constexpr double T16 = 0.19891236737965801;            // tan(pi/16)
constexpr double T316 = 0.66817863791929888;           // tan(3 pi/16)

// library calls for arguments outside the fast paths, kept out of line
static I2_HD __noinline__ double log_slow(double x) { return ::log(x); }
static I2_HD __noinline__ double atan2_slow(double y, double x) { return ::atan2(y, x); }

I2_HD __forceinline__ double log_pos(double x) {
    if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308)) return log_slow(x);
    int k;
    double m = frexp(x, &k);             // x = m 2^k, m in [0.5, 1)
    if (m < SQRT_HALF) { m += m; --k; }  // m in [sqrt(1/2), sqrt(2))
    const double f = m - 1.0;            // exact
    const double s = f / (m + 1.0);
    const double z = s * s;
    // 1/3 + z/5 + ... + z^8/19
    double p = 1.0 / 19.0;
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double s2 = s + s;
    const double r = fma(s2 * z, p, s2);          // 2 atanh(s)
    const double dk = (double)k;
    return fma(dk, LN2_HI, fma(dk, LN2_LO, r));
}

I2_HD __forceinline__ double atan2_fast(double y, double x) {
    const double ax = fabs(x), ay = fabs(y);
    if (!(ax <= 1.7976931348623157e308 && ay <= 1.7976931348623157e308) || (ax == 0.0 && ay == 0.0)) return atan2_slow(y, x);
    const bool swap = ay > ax;
    const double n = swap ? ax : ay;
    const double d = swap ? ay : ax;                 // t = n/d in [0, 1]
    const bool mid = n > T16 * d;
    const bool high = n > T316 * d;
    const double c = high ? 1.0 : (mid ? C8 : 0.0);
    const double bhi = high ? PIO4_HI : (mid ? ATAN_C8_HI : 0.0);
    const double blo = high ? PIO4_LO : (mid ? ATAN_C8_LO : 0.0);
    const double u = fma(-c, d, n) / fma(c, n, d);   // |u| <= tan(pi/16)
    const double w = u * u;
    // -1/3 + w/5 - ... + w^10/21
    double p = 1.0 / 21.0;
    p = fma(p, w, -1.0 / 19.0);
    p = fma(p, w, 1.0 / 17.0);
    p = fma(p, w, -1.0 / 15.0);
    p = fma(p, w, 1.0 / 13.0);
    p = fma(p, w, -1.0 / 11.0);
    p = fma(p, w, 1.0 / 9.0);
    p = fma(p, w, -1.0 / 7.0);
    p = fma(p, w, 1.0 / 5.0);
    p = fma(p, w, -1.0 / 3.0);
    const double au = fma(u * w, p, u);               // atan u
    double r = bhi + (au + blo);                      // atan t, t in [0, 1]
    if (swap) r = PIO2_HI - r + PIO2_LO;              // atan(1/t) = pi/2 - atan t
    if (x < 0.0) r = PI_HI - r + PI_LO;               // second and third quadrants
    return (y < 0.0 || (y == 0.0 && signbit(y))) ? -r : r;
}

}  // namespace i2m
