// exp, log and cbrt in FP64 from IEEE basic operations only (+, -, *, /,
// and the exact floor, frexp and ldexp), for the phase's leading term
// T0 f^(-5/3) - A. One source compiles for the host and for the device, and
// with contraction off both round every operation alike, so the host and
// the device get the same bits: the leading term reaches ~1e4 rad for a
// binary neutron star from 20 Hz, where one ulp of a library exp or log is
// ~1e-12 rad of phase, and the two platforms' libraries differ by that.
//
// Methods (textbook, e.g. J.-M. Muller, Elementary Functions, 3rd ed.,
// chs. 6 and 11):
//  * exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split
//    into a 21-bit head (k ln2_hi exact) and a tail; exp(r) by its Taylor
//    series to r^13/13! (truncation < 5e-18) in Horner form; 2^k by ldexp.
//  * log: x = 2^k m with m in [sqrt(1/2), sqrt(2)); f = m - 1 (exact), s =
//    f / (2 + f), log m = 2 atanh s = f - (f^2/2 - s (f^2/2 + R)), R =
//    sum_{j=1..11} 2 s^(2j) / (2j + 1) (|s| <= 0.172: truncation < 1e-19);
//    k ln2 added as head and tail.
//  * cbrt: exp(log|x| / 3) and one Newton step y - (y - x / y^2) / 3.
// Each is within about one ulp of the true value (tests/test_libm.py
// measures the error against long double).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

#if defined(__HIPCC__)
#define WFX_LIBM_HD __host__ __device__
#else
#define WFX_LIBM_HD
#endif
#if defined(__clang__)
#define WFX_LIBM_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define WFX_LIBM_NO_CONTRACT
#endif

namespace wfx_libm {

constexpr double kLn2Hi = 0.6931467056274414;       // ln 2 to 21 bits
constexpr double kLn2Lo = 4.7493250390316726e-07;   // ln 2 - kLn2Hi
constexpr double kInvLn2 = 1.4426950408889634;
constexpr double kSqrtHalf = 0.7071067811865476;

WFX_LIBM_HD inline double exp(double x) {
WFX_LIBM_NO_CONTRACT
    if (x != x) return x;
    if (x > 709.782712893384) return std::numeric_limits<double>::infinity();
    if (x < -745.1332191019412) return 0.0;
    const double k = floor(x * kInvLn2 + 0.5);
    const double hi = x - k * kLn2Hi;
    const double lo = k * kLn2Lo;
    const double r = hi - lo;
    // (exp(r) - 1 - r) / r^2 = sum_{j >= 2} r^(j-2) / j!
    double p = 1.6059043836821613e-10;
    p = p * r + 2.08767569878681e-09;
    p = p * r + 2.505210838544172e-08;
    p = p * r + 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    const double e = 1.0 + (hi - (lo - r * r * p));
    return ldexp(e, (int)k);
}

WFX_LIBM_HD inline double log(double x) {
WFX_LIBM_NO_CONTRACT
    if (!(x > 0.0)) return x == 0.0 ? -std::numeric_limits<double>::infinity()
                                    : std::numeric_limits<double>::quiet_NaN();
    if (x == std::numeric_limits<double>::infinity()) return x;
    int e = 0;
    double m = frexp(x, &e);  // x = m 2^e, m in [1/2, 1)
    if (m < kSqrtHalf) {
        m = m * 2.0;
        e -= 1;
    }
    const double k = (double)e;
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    // R = sum_{j >= 1} 2 z^j / (2j + 1)
    double q = 0.08695652173913043;
    q = q * z + 0.09523809523809523;
    q = q * z + 0.10526315789473684;
    q = q * z + 0.11764705882352941;
    q = q * z + 0.13333333333333333;
    q = q * z + 0.15384615384615385;
    q = q * z + 0.18181818181818182;
    q = q * z + 0.2222222222222222;
    q = q * z + 0.2857142857142857;
    q = q * z + 0.4;
    q = q * z + 0.6666666666666666;
    const double rr = z * q;
    const double hfsq = 0.5 * f * f;
    return k * kLn2Hi + (f - (hfsq - (s * (hfsq + rr) + k * kLn2Lo)));
}

WFX_LIBM_HD inline double cbrt(double x) {
WFX_LIBM_NO_CONTRACT
    if (x != x || x == 0.0 || x == std::numeric_limits<double>::infinity() ||
        x == -std::numeric_limits<double>::infinity())
        return x;
    const double a = x < 0.0 ? -x : x;
    double y = wfx_libm::exp(wfx_libm::log(a) / 3.0);
    y = y - (y - a / (y * y)) / 3.0;
    return x < 0.0 ? -y : y;
}

}  // namespace wfx_libm
