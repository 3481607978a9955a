// mcg_math.h -- portable FP64 logarithm and sin/cos(pi r) built only from
// IEEE-754 additions, multiplications and divisions.
//
// Compiled unchanged by g++ (host) and hipcc (gfx950), both with
// -ffp-contract=off, these produce bit-identical results on the CPU and the
// GPU.  The deterministic mode of the engine needs that: diffuse reflections
// between walls magnify a position offset by up to 1/cos(incidence) per bounce,
// so a one-ulp difference between two math libraries grows until two runs of
// the same bundle take different events after a few hundred bounces.
//
// Methods (textbook, no library constants):
//  * log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1,
//    s = f / (2 + f), log(1 + f) = 2 atanh(s) = 2 s sum_k s^2k / (2k + 1)
//    (the atanh series; |s| <= 0.1716 so 13 terms reach 1e-18), and
//    e ln 2 added with ln 2 split in two parts (Cody and Waite, "Software
//    Manual for the Elementary Functions", 1980) so that e * ln2_hi is exact.
//  * sin(pi r), cos(pi r) for |r| <= 1: r = j/2 + t with integer j and
//    |t| <= 1/4 (exact in binary), x = pi t, Taylor series of sin and cos
//    on |x| <= pi/4 to x^21 / x^22, then the quadrant j selects and signs.
// Accuracy (checked against glibc in tests/test_math.py): a few ulp.
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define MCG_HD __host__ __device__ __forceinline__
#else
#define MCG_HD inline
#endif

namespace mcg_math {

MCG_HD double bits_to_double(uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __longlong_as_double(static_cast<long long>(b));
#else
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
#endif
}

MCG_HD uint64_t double_to_bits(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    return static_cast<uint64_t>(__double_as_longlong(d));
#else
    uint64_t b;
    std::memcpy(&b, &d, sizeof b);
    return b;
#endif
}

// Natural logarithm of a positive normal double.
MCG_HD double log(double x) {
    const uint64_t b = double_to_bits(x);
    int e = int((b >> 52) & 0x7ff) - 1023;
    double m = bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);  // [1, 2)
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double f = m - 1.0;            // exact (Sterbenz)
    const double s = f / (2.0 + f);
    const double z = s * s;
    // sum_{k>=1} z^k / (2k + 1), Horner from the highest term
    double p = 1.0 / 27.0;
    p = 1.0 / 25.0 + z * p;
    p = 1.0 / 23.0 + z * p;
    p = 1.0 / 21.0 + z * p;
    p = 1.0 / 19.0 + z * p;
    p = 1.0 / 17.0 + z * p;
    p = 1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    const double two_s = 2.0 * s;
    const double log_m = two_s + two_s * (z * p);
    // ln 2 = ln2_hi + ln2_lo with ln2_hi holding 32 significant bits
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double de = double(e);
    return de * ln2_hi + (log_m + de * ln2_lo);
}

// sin(pi r) and cos(pi r) for |r| <= 1.
MCG_HD void sincospi(double r, double* sn, double* cs) {
    // nearest multiple of 1/2 (halves rounded away from zero)
    const double j = static_cast<double>(static_cast<int64_t>(r * 2.0 + (r < 0.0 ? -0.5 : 0.5)));
    const double t = r - 0.5 * j;   // |t| <= 1/4, exact
    const double x = t * 3.141592653589793;
    const double x2 = x * x;
    // sin x = x - x^3 (1/3! - x^2/5! + x^4/7! - ...), through x^21
    double ps = 1.0 / 51090942171709440000.0;  // 1/21!
    ps = -1.0 / 121645100408832000.0 + x2 * ps;  // 1/19!
    ps = 1.0 / 355687428096000.0 + x2 * ps;      // 1/17!
    ps = -1.0 / 1307674368000.0 + x2 * ps;       // 1/15!
    ps = 1.0 / 6227020800.0 + x2 * ps;           // 1/13!
    ps = -1.0 / 39916800.0 + x2 * ps;            // 1/11!
    ps = 1.0 / 362880.0 + x2 * ps;               // 1/9!
    ps = -1.0 / 5040.0 + x2 * ps;                // 1/7!
    ps = 1.0 / 120.0 + x2 * ps;                  // 1/5!
    ps = -1.0 / 6.0 + x2 * ps;                   // 1/3!
    const double sin_x = x + (x * x2) * ps;
    // cos x = 1 - x^2/2! + x^4/4! - ..., through x^22
    double pc = -1.0 / 1124000727777607680000.0;  // 1/22!
    pc = 1.0 / 2432902008176640000.0 + x2 * pc;    // 1/20!
    pc = -1.0 / 6402373705728000.0 + x2 * pc;      // 1/18!
    pc = 1.0 / 20922789888000.0 + x2 * pc;         // 1/16!
    pc = -1.0 / 87178291200.0 + x2 * pc;           // 1/14!
    pc = 1.0 / 479001600.0 + x2 * pc;              // 1/12!
    pc = -1.0 / 3628800.0 + x2 * pc;               // 1/10!
    pc = 1.0 / 40320.0 + x2 * pc;                  // 1/8!
    pc = -1.0 / 720.0 + x2 * pc;                   // 1/6!
    pc = 1.0 / 24.0 + x2 * pc;                     // 1/4!
    pc = -0.5 + x2 * pc;                           // 1/2!
    const double cos_x = 1.0 + x2 * pc;
    // sin(pi (t + j/2)): the quarter-turn count j mod 4 rotates (sin, cos)
    const int q = static_cast<int>(j) & 3;
    if (q == 0) {
        *sn = sin_x;
        *cs = cos_x;
    } else if (q == 1) {
        *sn = cos_x;
        *cs = -sin_x;
    } else if (q == 2) {
        *sn = -sin_x;
        *cs = -cos_x;
    } else {
        *sn = -cos_x;
        *cs = sin_x;
    }
}

}  // namespace mcg_math
