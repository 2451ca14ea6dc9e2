/* Oracle exp and log, FP64, from published methods (no libm):
 *
 *   exp: Cody & Waite argument reduction (W. J. Cody, W. Waite, "Software
 *        Manual for the Elementary Functions", 1980): x = k ln2 + r with
 *        ln2 split into a high part exact under multiplication by k and a
 *        low part, |r| <= ln2/2, then the Taylor series of e^r to degree 13
 *        by Horner (truncation < 4e-18 relative), scaled by 2^k through the
 *        exponent bits.
 *   log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *        s = (m - 1)/(m + 1), |s| <= 0.1716, atanh by its odd series to
 *        s^23 (truncation < 1e-18 relative), then e ln2 added in two parts.
 *
 * The GPU kernels restate the same two routines operation for operation,
 * and every file that computes device physics on either side is compiled
 * without floating-point contraction, so device evaluation is bitwise equal
 * between this oracle and the HIP lanes. The reference's own exp/log live
 * in Gompute (gompute.math, src/device/eval.zig:41-46), which is not in
 * the reference tree: their bits are parity unpinned. These are pinned
 * against glibc to a few ulp in tests/test_oracle_math.py.
 */
#include <stdint.h>
#include <string.h>
#include "orc.h"

static double from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* ln2 split: 0x3FE62E42FEE00000 has 11 trailing zero mantissa bits. */
#define LN2_HI 6.93147180369123816490e-01
#define LN2_LO 1.90821492927058770002e-10
#define LOG2E 1.44269504088896338700e+00

double orc_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893383973096) return from_bits(0x7ff0000000000000ull);
    if (x < -745.133219101941108420) return 0.0;
    double t = x * LOG2E;
    /* Round half away from zero through a truncating conversion. */
    int k = (int)(t + (t >= 0.0 ? 0.5 : -0.5));
    double kd = (double)k;
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double p = 1.0 / 6227020800.0;              /* 1/13! */
    p = 1.0 / 479001600.0 + r * p;              /* 1/12! */
    p = 1.0 / 39916800.0 + r * p;               /* 1/11! */
    p = 1.0 / 3628800.0 + r * p;                /* 1/10! */
    p = 1.0 / 362880.0 + r * p;                 /* 1/9! */
    p = 1.0 / 40320.0 + r * p;                  /* 1/8! */
    p = 1.0 / 5040.0 + r * p;                   /* 1/7! */
    p = 1.0 / 720.0 + r * p;                    /* 1/6! */
    p = 1.0 / 120.0 + r * p;                    /* 1/5! */
    p = 1.0 / 24.0 + r * p;                     /* 1/4! */
    p = 1.0 / 6.0 + r * p;                      /* 1/3! */
    p = 0.5 + r * p;                            /* 1/2! */
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    /* p * 2^k, in two steps below the normal range. */
    if (k > 1023) {
        p = p * 2.0;
        k -= 1;
    }
    if (k < -1021) {
        double s1 = from_bits((uint64_t)(k + 54 + 1023) << 52);
        return (p * s1) * from_bits((uint64_t)(-54 + 1023) << 52);
    }
    return p * from_bits((uint64_t)(k + 1023) << 52);
}

#define SQRT_HALF 7.07106781186547524401e-01

double orc_log(double x)
{
    uint64_t b = to_bits(x);
    if (x != x) return x;
    if (x < 0.0) return from_bits(0x7ff8000000000000ull);
    if (x == 0.0) return from_bits(0xfff0000000000000ull);
    if (b == 0x7ff0000000000000ull) return x;
    int e = 0;
    if (b < 0x0010000000000000ull) {            /* subnormal */
        x = x * 18014398509481984.0;             /* 2^54 */
        b = to_bits(x);
        e = -54;
    }
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.0 / SQRT_HALF) {
        m = m * 0.5;
        e += 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
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
    double lm = 2.0 * s + (2.0 * s) * (s2 * p);
    double ed = (double)e;
    return ed * LN2_HI + (ed * LN2_LO + lm);
}
