/*
 * ngmath.h -- double-precision exp, log and pow shared by the HIP kernels and
 * the CPU oracle, so that both evaluate every transcendental bit for bit alike.
 *
 * Only IEEE-754 operations that are correctly rounded on both x86-64 and CDNA
 * are used (+ - * /, fma, rint, frexp, ldexp); compile with -ffp-contract=off
 * so that no other multiply-add is fused on either side.
 *
 * Methods (published; no vendor library code is reproduced):
 *  - exp: Cody & Waite argument reduction x = k*ln2 + r with ln2 split into a
 *    high part with trailing zero bits and a low part (W. J. Cody, W. Waite,
 *    "Software Manual for the Elementary Functions", 1980; the split values
 *    NG_LN2_HI / NG_LN2_LO and the log reduction below are the published
 *    ones of Sun's freely distributable fdlibm, e_exp.c / e_log.c), then the Taylor
 *    series of e^r to degree 13 by Horner's rule (|r| <= ln2/2 makes the
 *    truncation error < 2^-57), then scaling by 2^k.
 *  - log: x = 2^k * m with m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f/(2+f),
 *    log(1+f) = 2 atanh(s) = f - f^2/2 + s*(f^2/2 + R(s^2)) where
 *    R(z) = sum_{j>=1} 2 z^j / (2j+1) is the atanh series (Cody & Waite, ibid.),
 *    summed to j = 11; the k*ln2 term uses the same split constant.
 *  - pow: small integer exponents by repeated multiplication, otherwise
 *    exp(y * log(x)).
 * Accuracy against glibc is checked in tests/test_oracle.py
 * (test_shared_exp_log_within_one_ulp_of_glibc).
 */
#ifndef NGMATH_H
#define NGMATH_H

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define NG_HD __host__ __device__ static inline
#else
#include <math.h>
#define NG_HD static inline
#endif

#define NG_LN2_HI 6.93147180369123816490e-01   /* ln 2 rounded to 32 significant bits */
#define NG_LN2_LO 1.90821492927058770002e-10   /* ln 2 - NG_LN2_HI                     */
#define NG_LOG2E 1.44269504088896338700e+00
#define NG_SQRT_HALF 0.70710678118654752440

NG_HD double ng_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893383973096) return 1.0 / 0.0;
    if (x < -745.133219101941108420) return 0.0;
    double kf = rint(x * NG_LOG2E);
    double r = (x - kf * NG_LN2_HI) - kf * NG_LN2_LO;
    /* Taylor coefficients 1/n! for n = 13 .. 2 */
    double p = 1.6059043836821614599e-10;          /* 1/13! */
    p = fma(p, r, 2.0876756987868098979e-09);       /* 1/12! */
    p = fma(p, r, 2.5052108385441718775e-08);       /* 1/11! */
    p = fma(p, r, 2.7557319223985890653e-07);       /* 1/10! */
    p = fma(p, r, 2.7557319223985890653e-06);       /* 1/9!  */
    p = fma(p, r, 2.4801587301587301566e-05);       /* 1/8!  */
    p = fma(p, r, 1.9841269841269841253e-04);       /* 1/7!  */
    p = fma(p, r, 1.3888888888888888889e-03);       /* 1/6!  */
    p = fma(p, r, 8.3333333333333333333e-03);       /* 1/5!  */
    p = fma(p, r, 4.1666666666666666667e-02);       /* 1/4!  */
    p = fma(p, r, 1.6666666666666666667e-01);       /* 1/3!  */
    p = fma(p, r, 0.5);                             /* 1/2!  */
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)kf);
}

NG_HD double ng_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return 0.0 / 0.0;
    if (x == 0.0) return -1.0 / 0.0;
    if (x == 1.0 / 0.0) return x;
    int k;
    double m = frexp(x, &k);                   /* x = m * 2^k, m in [0.5, 1) */
    if (m < NG_SQRT_HALF) { m = m * 2.0; k = k - 1; }
    double f = m - 1.0;                        /* exact */
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 0.0869565217391304347826;       /* 2/23 */
    R = fma(R, z, 0.0952380952380952380952);   /* 2/21 */
    R = fma(R, z, 0.1052631578947368421053);   /* 2/19 */
    R = fma(R, z, 0.1176470588235294117647);   /* 2/17 */
    R = fma(R, z, 0.1333333333333333333333);   /* 2/15 */
    R = fma(R, z, 0.1538461538461538461538);   /* 2/13 */
    R = fma(R, z, 0.1818181818181818181818);   /* 2/11 */
    R = fma(R, z, 0.2222222222222222222222);   /* 2/9  */
    R = fma(R, z, 0.2857142857142857142857);   /* 2/7  */
    R = fma(R, z, 0.4);                        /* 2/5  */
    R = fma(R, z, 0.6666666666666666666667);   /* 2/3  */
    R = R * z;
    double hfsq = 0.5 * f * f;
    double kd = (double)k;
    return kd * NG_LN2_HI + (f - (hfsq - (s * (hfsq + R) + kd * NG_LN2_LO)));
}

NG_HD double ng_pow(double x, double y)
{
    if (y == 1.0) return x;
    if (y == 2.0) return x * x;
    if (y == 3.0) return x * x * x;
    if (y == 4.0) { double x2 = x * x; return x2 * x2; }
    if (y == 0.0 || x == 1.0) return 1.0;
    if (x == 0.0) return y > 0.0 ? 0.0 : 1.0 / 0.0;
    if (x < 0.0) return 0.0 / 0.0;
    return ng_exp(y * ng_log(x));
}

#endif
