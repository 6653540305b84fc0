// i2_math.hpp -- FP64 square root, logarithm and two-argument arctangent
// for the quadrature hot loop.
//
// The library versions cost about 16 (sqrt), 92 (log) and 99 (atan2) VALU
// instructions on gfx950; the inner integral needs three of each of the
// first two and one atan2 per quadrature point, so they dominate the kernel.
// These versions drop the range handling the hot loop never needs, are
// branch-free on the normal path, and stay within about 2 ulp (checked on the
// host against 120-bit references).
//
// Methods (textbook):
//   sqrt:  s > 0 normal: y ~ 1/sqrt(s) from the hardware estimate, one
//          Goldschmidt step g = s y, h = y/2, r = 1/2 - g h, then
//          g + (s - g^2) h (one correction; within 1 ulp, not always
//          correctly rounded).
//   log:   x = 2^k m with m in [sqrt(1/2), sqrt(2)); j = rint(64 (m - 1)) picks
//          c_j = 1 + j/64 from a 47-entry table of invc_j = 1/c_j (rounded) and
//          T_j = -log(invc_j) (hi + lo); r = m invc_j - 1 by one fma, |r| < 0.0112;
//          log1p(r) by its Taylor series to r^8 (truncation below 2^-56
//          relative); log x = (k ln2_hi + T_hi) + (k ln2_lo + T_lo + log1p r).
//          c_0 = 1 exactly, so there is no cancellation near x = 1.  No division.
//   atan2: reduce to t = min(|x|,|y|)/max(|x|,|y|) in [0, 1]; with a centre
//          c in {0, c8, 1} (c8 = the double nearest tan(pi/8)),
//          atan t = atan c + atan u, u = (n - c d)/(d + c n) for t = n/d, so
//          |u| <= tan(pi/16); atan u by its Taylor series to u^21/21;
//          then the octant and quadrant identities.
//   1/x and a/b (rcp_pos, div_pos): hardware reciprocal estimate refined by
//          Newton steps with fma residuals; used where the operand is a
//          product of distances, far inside [2^-1000, 2^1000].
// Non-finite, zero, negative and subnormal arguments take the library path
// (log, atan2).
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

// estimate of 1/sqrt(s): the hardware instruction on the device, the library on the host
__device__ __forceinline__ double rsq_estimate(double s) { return __builtin_amdgcn_rsq(s); }
__host__ inline double rsq_estimate(double s) { return 1.0 / ::sqrt(s); }

// estimate of 1/x: the hardware instruction on the device; on the host a
// single-precision reciprocal of the mantissa, so that host checks exercise
// the refinement from a low-precision start
__device__ __forceinline__ double rcp_estimate(double x) { return __builtin_amdgcn_rcp(x); }
__host__ inline double rcp_estimate(double x) {
    int k;
    const double m = ::frexp(x, &k);
    return ::ldexp((double)(1.0f / (float)m), -k);
}

// 1/x for |x| in [2^-1000, 2^1000]: estimate, then three Newton steps
// y += y (1 - x y), each with the residual exact in one fma
I2_HD __forceinline__ double rcp_pos(double x) {
    double y = rcp_estimate(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// a / b for |b| in [2^-1000, 2^1000]: two Newton steps on 1/b, then the
// quotient with one residual correction
I2_HD __forceinline__ double div_pos(double a, double b) {
    double y = rcp_estimate(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    const double q = a * y;
    return fma(fma(-b, q, a), y, q);
}

// sqrt of a positive normal number (squared distances of the hot loop)
I2_HD __forceinline__ double sqrt_pos(double s) {
    const double y = rsq_estimate(s);
    double g = s * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, s);
    return fma(d, h, g);
}

// a * b + c for a Horner step whose addend c is a constant.  On the device
// this is one three-operand v_fma_f64: left to itself the compiler keeps such
// constants in VGPRs and emits a copy into the accumulator before each
// two-operand v_fmac_f64.
I2_HD __forceinline__ double fma_k(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
#else
    return fma(a, b, c);
#endif
}

// log reduction table (gen_log_table.py): entry j + LOGTAB_J0 for c_j = 1 + j/64,
// held as three arrays so that an entry's address is a shift of j
struct LogEntry { double invc, thi, tlo; };
struct LogTable { double invc[47], thi[47], tlo[47]; };
constexpr int LOGTAB_J0 = 19;
constexpr int LOGTAB_N = 47;
constexpr LogEntry h_logent[LOGTAB_N] = {
#include "i2_logtab.inc"
};
constexpr LogTable make_logtab(const LogEntry (&e)[LOGTAB_N]) {
    LogTable t{};
    for (int k = 0; k < LOGTAB_N; ++k) { t.invc[k] = e[k].invc; t.thi[k] = e[k].thi; t.tlo[k] = e[k].tlo; }
    return t;
}
static constexpr LogTable h_logtab = make_logtab(h_logent);
static __constant__ LogEntry c_logtab[LOGTAB_N] = {
#include "i2_logtab.inc"
};
// copies the table into `s` (LDS) with the threads of the block; the caller synchronises
__device__ __forceinline__ void fill_logtab(LogTable& s) {
    for (int k = threadIdx.x; k < LOGTAB_N; k += blockDim.x) {
        s.invc[k] = c_logtab[k].invc; s.thi[k] = c_logtab[k].thi; s.tlo[k] = c_logtab[k].tlo;
    }
}

// log of a positive normal number with the reduction table `tab` (LDS on the device)
I2_HD __forceinline__ double log_tab(double x, const LogTable* __restrict__ tab) {
    if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308)) return log_slow(x);
    int k;
    double m = frexp(x, &k);             // x = m 2^k, m in [0.5, 1)
    if (m < SQRT_HALF) { m += m; --k; }  // m in [sqrt(1/2), sqrt(2))
    // nearest centre (ties upward): 64 m + 1/2 in [45.7, 91.1), so j in [-19, 27]
    const int j = (int)fma(m, 64.0, 0.5) - 64 + LOGTAB_J0;
    const double r = fma(m, tab->invc[j], -1.0);
    // log1p(r) = r - r^2/2 + ... - r^8/8
    double q = fma_k(r, -1.0 / 8.0, 1.0 / 7.0);
    q = fma_k(q, r, -1.0 / 6.0);
    q = fma_k(q, r, 1.0 / 5.0);
    q = fma(q, r, -1.0 / 4.0);
    q = fma_k(q, r, 1.0 / 3.0);
    q = fma(q, r, -0.5);
    const double p = fma(r * r, q, r);
    const double dk = (double)k;
    return fma(dk, LN2_HI, tab->thi[j]) + (fma(dk, LN2_LO, tab->tlo[j]) + p);
}

I2_HD __forceinline__ double atan2_fast(double y, double x) {
    const double ax = fabs(x), ay = fabs(y);
    // zero, tiny, huge, infinite and NaN arguments take the library path
    const double sum = ax + ay;
    if (!(sum >= 0x1p-999 && sum <= 0x1p999)) return atan2_slow(y, x);
    const bool swap = ay > ax;
    const double n = fmin(ax, ay);
    const double d = fmax(ax, ay);                   // t = n/d in [0, 1], d >= 2^-1000
    const bool mid = n > T16 * d;
    const bool high = n > T316 * d;
    const double c = high ? 1.0 : (mid ? C8 : 0.0);
    const double bhi = high ? PIO4_HI : (mid ? ATAN_C8_HI : 0.0);
    const double blo = high ? PIO4_LO : (mid ? ATAN_C8_LO : 0.0);
    const double u = div_pos(fma(-c, d, n), fma(c, n, d));   // |u| <= tan(pi/16)
    const double w = u * u;
    // -1/3 + w/5 - ... + w^10/21
    double p = fma_k(w, 1.0 / 21.0, -1.0 / 19.0);
    p = fma_k(p, w, 1.0 / 17.0);
    p = fma_k(p, w, -1.0 / 15.0);
    p = fma_k(p, w, 1.0 / 13.0);
    p = fma_k(p, w, -1.0 / 11.0);
    p = fma_k(p, w, 1.0 / 9.0);
    p = fma_k(p, w, -1.0 / 7.0);
    p = fma_k(p, w, 1.0 / 5.0);
    p = fma_k(p, w, -1.0 / 3.0);
    const double au = fma(u * w, p, u);               // atan u
    double r = bhi + (au + blo);                      // atan t, t in [0, 1]
    if (swap) r = PIO2_HI - r + PIO2_LO;              // atan(1/t) = pi/2 - atan t
    if (x < 0.0) r = PI_HI - r + PI_LO;               // second and third quadrants
    return (y < 0.0 || (y == 0.0 && signbit(y))) ? -r : r;
}

}  // namespace i2m
