// exp and log in FP64 for the lane engine, host and device.
//
// exp: Cody & Waite argument reduction (W. J. Cody, W. Waite, "Software
// Manual for the Elementary Functions", 1980): x = k ln2 + r, |r| <= ln2/2,
// with ln2 split so k * LN2_HI is exact; e^r by its Taylor series to degree 13
// in Horner form; the result scaled by 2^k through the exponent field.
// log: x = 2^e m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s) with
// s = (m - 1)/(m + 1), the odd atanh series to s^23; e ln2 added in two parts.
//
// Every caller compiles without floating-point contraction
// (-ffp-contract=off), so host and device return the same bits for the same
// input. Error against a correctly rounded result is a few ulp.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define ESL_HD __host__ __device__ inline
#else
#define ESL_HD inline
#endif

namespace esl {

ESL_HD double bits_to_double(uint64_t b) { return __builtin_bit_cast(double, b); }
ESL_HD uint64_t double_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kLog2e = 1.44269504088896338700e+00;
constexpr double kSqrtHalf = 7.07106781186547524401e-01;

ESL_HD double exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893383973096) return bits_to_double(0x7ff0000000000000ull);
    if (x < -745.133219101941108420) return 0.0;
    const double t = x * kLog2e;
    int k = (int)(t + (t >= 0.0 ? 0.5 : -0.5));
    const double kd = (double)k;
    const double r = (x - kd * kLn2Hi) - kd * kLn2Lo;
    double p = 1.0 / 6227020800.0;
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    if (k > 1023) {
        p = p * 2.0;
        k -= 1;
    }
    if (k < -1021) {
        const double s1 = bits_to_double((uint64_t)(k + 54 + 1023) << 52);
        return (p * s1) * bits_to_double((uint64_t)(-54 + 1023) << 52);
    }
    return p * bits_to_double((uint64_t)(k + 1023) << 52);
}

ESL_HD double log(double x)
{
    uint64_t b = double_to_bits(x);
    if (x != x) return x;
    if (x < 0.0) return bits_to_double(0x7ff8000000000000ull);
    if (x == 0.0) return bits_to_double(0xfff0000000000000ull);
    if (b == 0x7ff0000000000000ull) return x;
    int e = 0;
    if (b < 0x0010000000000000ull) {
        x = x * 18014398509481984.0;
        b = double_to_bits(x);
        e = -54;
    }
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = bits_to_double((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.0 / kSqrtHalf) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = 1.0 / 21.0 + s2 * p;
    p = 1.0 / 19.0 + s2 * p;
    p = 1.0 / 17.0 + s2 * p;
    p = 1.0 / 15.0 + s2 * p;
    p = 1.0 / 13.0 + s2 * p;
    p = 1.0 / 11.0 + s2 * p;
    p = 1.0 / 9.0 + s2 * p;
    p = 1.0 / 7.0 + s2 * p;
    p = 1.0 / 5.0 + s2 * p;
    p = 1.0 / 3.0 + s2 * p;
    const double lm = 2.0 * s + (2.0 * s) * (s2 * p);
    const double ed = (double)e;
    return ed * kLn2Hi + (ed * kLn2Lo + lm);
}

}  // namespace esl
