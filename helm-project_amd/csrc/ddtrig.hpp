// ddtrig.hpp -- device sin and cos of a double to ~100 bits (double-double
// arithmetic), and where the correctly rounded double lies relative to the
// rounding midpoint next to it.
//
// Use: node xyz of a mesh on the device with host libm's bits.  glibc's
// sin/cos/sincos (2.28 and later) return results within ~0.55 ULP of the
// exact value, so when the exact value is more than 0.05 ULP away from the
// midpoint between the two doubles around it, libm returns the correctly
// rounded double, which these functions compute.  Values closer to a midpoint
// than kLibmBand ULP are reported; for them the caller asks host libm, whose
// answer is one of the two doubles around the exact value.  Results below
// kTinyResult (arguments near a multiple of pi/2, where libm's error relative
// to the result grows) are left to host libm altogether.
//
// Method: Cody-Waite reduction by a three-double pi/2 (x - n pi/2 carried as a
// double-double; the first step is exact by Sterbenz's lemma, the products
// by fma), then Taylor series of sin and cos on |r| <= pi/4 (+ rounding slack)
// in Horner form over double-double coefficients 1/k!, to k = 23 / 22: the
// truncation error is below 2^-86 relative.  The constants are pi/2 and 1/k!
// rounded to three and two doubles.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace axis_hip::detail::dd {

/// Distance (ULP) from a rounding midpoint below which host libm's answer is
/// asked for: above glibc's 0.55 - 0.5 ULP with a margin.
constexpr double kLibmBand = 0.0625;
/// |x| beyond this (radians) is left to host libm (n pi/2 reduction bound).
constexpr double kMaxArg = 64.0;
/// Results smaller than this in magnitude (x within ~0.06 degrees of a
/// multiple of pi/2) are left to host libm: there libm's error relative to
/// the tiny result is larger than elsewhere (0.56 ULP seen one double above
/// pi/2).
constexpr double kTinyResult = 0x1p-10;

struct DD {
    double hi, lo;
};

__host__ __device__ inline DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}

__host__ __device__ inline DD quick_two_sum(double a, double b) {  // |a| >= |b|
    const double s = a + b;
    return DD{s, b - (s - a)};
}

__host__ __device__ inline DD two_prod(double a, double b) {
    const double p = a * b;
    return DD{p, fma(a, b, -p)};
}

__host__ __device__ inline DD add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    const DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

__host__ __device__ inline DD add_d(DD a, double b) {
    DD s = two_sum(a.hi, b);
    s.lo += a.lo;
    return quick_two_sum(s.hi, s.lo);
}

__host__ __device__ inline DD mul(DD a, DD b) {
    DD p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p.hi, p.lo);
}

/// sin(x) and cos(x) as double-doubles (hi the correctly rounded double up to
/// an exact tie); false when |x| > kMaxArg or x is not finite.
__host__ __device__ inline bool sincos_dd(double x, DD &s, DD &c) {
    // pi/2 = kPio2[0] + kPio2[1] + kPio2[2] (+ O(2^-163))
    constexpr double kPio2[3] = {0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54, -0x1.f1976b7ed8fbcp-110};
    constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;

    // (-1)^k / (2k+1)! for k = 1..11 and (-1)^k / (2k)! for k = 1..11, as (hi, lo)
    constexpr double kSinC[11][2] = {
        {-0x1.5555555555555p-3, -0x1.5555555555555p-57}, {0x1.1111111111111p-7, 0x1.1111111111111p-63},
        {-0x1.a01a01a01a01ap-13, -0x1.a01a01a01a01ap-73}, {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
        {-0x1.ae64567f544e4p-26, 0x1.c062e06d1f209p-80}, {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},
        {-0x1.ae7f3e733b81fp-41, -0x1.1d8656b0ee8cbp-97}, {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},
        {-0x1.2f49b46814157p-57, -0x1.2650f61dbdcb4p-112}, {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},
        {-0x1.761b41316381ap-75, 0x1.3423c7d91404fp-130}};
    constexpr double kCosC[11][2] = {
        {-0x1.0000000000000p-1, 0.0}, {0x1.5555555555555p-5, 0x1.5555555555555p-59},
        {-0x1.6c16c16c16c17p-10, 0x1.f49f49f49f49fp-65}, {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
        {-0x1.27e4fb7789f5cp-22, -0x1.cbbc05b4fa99ap-76}, {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
        {-0x1.93974a8c07c9dp-37, -0x1.05d6f8a2efd1fp-92}, {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101},
        {-0x1.6827863b97d97p-53, -0x1.eec01221a8b0bp-107}, {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120},
        {-0x1.0ce396db7f853p-70, 0x1.aebcdbd20331cp-124}};
    if (!(fabs(x) <= kMaxArg)) return false;
    const double n = rint(x * kTwoOverPi);
    DD r{x, 0.0};
    if (n != 0.0) {
        const DD p1 = two_prod(n, kPio2[0]);
        const DD p2 = two_prod(n, kPio2[1]);
        r = two_sum(x - p1.hi, -p1.lo);  // x - p1.hi is exact: p1.hi / 2 <= x <= 2 p1.hi
        r = add_d(r, -p2.hi);
        r = add_d(r, -p2.lo);
        r = add_d(r, -(n * kPio2[2]));
    }
    const DD z = mul(r, r);
    DD ps{kSinC[10][0], kSinC[10][1]}, pc{kCosC[10][0], kCosC[10][1]};
#pragma unroll
    for (int k = 9; k >= 0; --k) {
        ps = add(mul(ps, z), DD{kSinC[k][0], kSinC[k][1]});
        pc = add(mul(pc, z), DD{kCosC[k][0], kCosC[k][1]});
    }
    const DD sr = add(r, mul(mul(r, z), ps));   // r + r^3 (-1/3! + ...)
    const DD cr = add_d(mul(z, pc), 1.0);       // 1 + r^2 (-1/2! + ...)
    const int q = static_cast<int>(static_cast<long long>(n) & 3);
    const DD ns{-sr.hi, -sr.lo}, nc{-cr.hi, -cr.lo};
    s = q == 0 ? sr : q == 1 ? cr : q == 2 ? ns : nc;
    c = q == 0 ? cr : q == 1 ? ns : q == 2 ? nc : sr;
    return true;
}

/// Distance (ULP) of the value hi + lo from the midpoint between the doubles
/// around it (0.5: exactly a double; 0: a tie).  hi + lo normalised, hi =
/// round(hi + lo).
__host__ __device__ inline double midpoint_distance(DD v) {
    if (v.lo == 0.0) return 0.5;
    int e = 0;
    const double m = frexp(fabs(v.hi), &e);  // |hi| = m 2^e, m in [0.5, 1)
    const double up = ldexp(1.0, e - 53);    // spacing above |hi|
    const bool toward_zero = (v.lo < 0.0) != (v.hi < 0.0);
    const double sp = (toward_zero && m == 0.5) ? 0.5 * up : up;
    return 0.5 - fabs(v.lo) / sp;
}

/// Whether host libm may return other than v.hi: v near a rounding midpoint.
__host__ __device__ inline bool near_midpoint(DD v) { return midpoint_distance(v) < kLibmBand; }

/// Whether v is left to host libm entirely (tiny result).
__host__ __device__ inline bool tiny(DD v) { return fabs(v.hi) < kTinyResult; }

/// The double that host libm returns for v: hi, or (libm_lsb disagreeing with
/// hi's last bit) the double next to hi on the side of the exact value.
__host__ __device__ inline double pick(DD v, unsigned libm_lsb) {
    const long long b = __builtin_bit_cast(long long, v.hi);
    if (static_cast<unsigned>(b & 1) == libm_lsb) return v.hi;
    // one step in magnitude: away from zero when lo has hi's sign
    const bool away = (v.lo < 0.0) == (v.hi < 0.0);
    return __builtin_bit_cast(double, away ? b + 1 : b - 1);
}

}  // namespace axis_hip::detail::dd
