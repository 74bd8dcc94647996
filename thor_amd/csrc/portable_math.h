// Portable elementary functions built only from IEEE-754 +, -, *, / and
// exponent-bit manipulation, so host and device produce identical bits when
// compiled without FMA contraction. Used by the parity build
// (-DTHOR_PORTABLE_MATH) of both the HIP kernels and the CPU oracle, which
// removes libm differences (glibc vs ocml) from the comparison; production
// builds call the platform's pow/log/sin/cos.
//
// Accuracy (checked against glibc in tests/test_oracle.py, test_elementary_functions
// and test_portable_pow_over_the_step_ranges): log <= 2
// ulp, exp <= 1 ulp, pow <= 2.5 (1 + |y log x|) ulp (x > 0; the error of log x is
// multiplied by |y log x| in the exponent), sin/cos <= 1e-15
// absolute for |x| <= 4; not a replacement for libm elsewhere.
#pragma once

#ifdef __HIPCC__
#define PM_FN __host__ __device__ inline
#else
#define PM_FN inline
#endif

#include <cstdint>
#include <cmath>

namespace pm {

PM_FN uint64_t bits(double x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
}
PM_FN double from_bits(uint64_t u) {
    double x;
    __builtin_memcpy(&x, &u, 8);
    return x;
}

constexpr double LN2_HI = 6.93147180369123816490e-01;  // high 32 bits of ln 2
constexpr double LN2_LO = 1.90821492927058770002e-10;  // ln 2 - LN2_HI
constexpr double INV_LN2 = 1.44269504088896338700e+00;

// Natural logarithm of x > 0: x = m * 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m - 1) / (m + 1) (|s| <= 0.172), by the odd power
// series of atanh in Horner form (terms to s^27).
PM_FN double log(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -1.0 / 0.0 : 0.0 / 0.0;  // log 0 = -inf; negative or NaN -> NaN
    if (x == 1.0 / 0.0) return x;
    int scale = 0;
    if (x < 2.2250738585072014e-308) {  // subnormal: normalise
        x *= 18014398509481984.0;        // 2^54
        scale = -54;
    }
    uint64_t u = bits(x);
    int e = (int)((u >> 52) & 0x7ff) - 1023 + scale;
    double m = from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    // 1/(2k+1), k = 13..1
    double q = 1.0 / 27.0;
    q = q * s2 + 1.0 / 25.0;
    q = q * s2 + 1.0 / 23.0;
    q = q * s2 + 1.0 / 21.0;
    q = q * s2 + 1.0 / 19.0;
    q = q * s2 + 1.0 / 17.0;
    q = q * s2 + 1.0 / 15.0;
    q = q * s2 + 1.0 / 13.0;
    q = q * s2 + 1.0 / 11.0;
    q = q * s2 + 1.0 / 9.0;
    q = q * s2 + 1.0 / 7.0;
    q = q * s2 + 1.0 / 5.0;
    q = q * s2 + 1.0 / 3.0;
    const double lm = 2.0 * s + 2.0 * s * (s2 * q);
    const double de = (double)e;
    return (de * LN2_HI + lm) + de * LN2_LO;
}

// e^x for |x| < 700: x = n ln2 + r, |r| <= ln2 / 2, e^r by its Taylor series
// in Horner form (terms to r^17).
PM_FN double exp(double x) {
    if (!(x == x)) return x;                            // NaN
    if (x > 709.782712893384) return 1.0 / 0.0;         // overflow
    if (x < -745.1332191019412) return 0.0;             // below the smallest subnormal
    const double t = x * INV_LN2;
    const double n = (double)(int64_t)(t < 0 ? t - 0.5 : t + 0.5);
    const double r = (x - n * LN2_HI) - n * LN2_LO;
    double q = 1.0 / 355687428096000.0;   // 1/17!
    q = q * r + 1.0 / 20922789888000.0;   // 1/16!
    q = q * r + 1.0 / 1307674368000.0;    // 1/15!
    q = q * r + 1.0 / 87178291200.0;      // 1/14!
    q = q * r + 1.0 / 6227020800.0;       // 1/13!
    q = q * r + 1.0 / 479001600.0;        // 1/12!
    q = q * r + 1.0 / 39916800.0;         // 1/11!
    q = q * r + 1.0 / 3628800.0;          // 1/10!
    q = q * r + 1.0 / 362880.0;           // 1/9!
    q = q * r + 1.0 / 40320.0;            // 1/8!
    q = q * r + 1.0 / 5040.0;             // 1/7!
    q = q * r + 1.0 / 720.0;              // 1/6!
    q = q * r + 1.0 / 120.0;              // 1/5!
    q = q * r + 1.0 / 24.0;               // 1/4!
    q = q * r + 1.0 / 6.0;                // 1/3!
    q = q * r + 0.5;
    const double sum = 1.0 + (r + r * (r * q));
    const int64_t ni = (int64_t)n;
    if (ni > 1023)  // near the largest double: 2^ni is not representable, scale in two steps
        return (sum * from_bits((uint64_t)(ni - 1 + 1023) << 52)) * 2.0;
    if (ni < -1021)  // subnormal result: scale in two steps so both factors are normal
        return (sum * from_bits((uint64_t)(ni + 54 + 1023) << 52)) * 5.551115123125783e-17;  // 2^-54
    return sum * from_bits((uint64_t)(ni + 1023) << 52);
}

// x^y for x > 0; y = 0 and 1 exactly (as pow does), small integer exponents
// by repeated multiplication.
PM_FN double pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (y == 1.0) return x;
    if (y == 2.0) return x * x;
    if (y == 3.0) return x * x * x;
    if (y == 4.0) {
        const double x2 = x * x;
        return x2 * x2;
    }
    return exp(y * log(x));
}

// Reduces x to [-pi, pi] (x unchanged when already there).
PM_FN double reduce_2pi(double x) {
    const double twopi = 6.283185307179586;
    const double n = std::rint(x / twopi);
    return n == 0.0 ? x : x - n * twopi;
}

// sin and cos by their Taylor series after reduction to [-pi, pi].
PM_FN double sin(double x) {
    x = reduce_2pi(x);
    const double x2 = x * x;
    double term = x, sum = x;
    for (int k = 1; k <= 14; k++) {
        term = -term * x2 / (double)((2 * k) * (2 * k + 1));
        sum = sum + term;
    }
    return sum;
}
PM_FN double cos(double x) {
    x = reduce_2pi(x);
    const double x2 = x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k <= 14; k++) {
        term = -term * x2 / (double)((2 * k - 1) * (2 * k));
        sum = sum + term;
    }
    return sum;
}

} // namespace pm

// Elementary functions of the dynamical core: portable in the parity build,
// libm otherwise.
#ifdef THOR_PORTABLE_MATH
#define MPOW(x, y) pm::pow((x), (y))
#define MLOG(x) pm::log(x)
#define MSIN(x) pm::sin(x)
#define MCOS(x) pm::cos(x)
#define MEXP(x) pm::exp(x)
#else
#define MPOW(x, y) pow((x), (y))
#define MLOG(x) log(x)
#define MSIN(x) sin(x)
#define MCOS(x) cos(x)
#define MEXP(x) exp(x)
#endif
