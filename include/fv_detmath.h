/*
 * fv_detmath.h -- deterministic double-precision elementary functions used by
 * the time step wherever the reference calls a transcendental intrinsic:
 *
 *   TANH  RAMP          (internal_step.F:115-120, external_step.F:83-88)
 *   DCOS  spectral tide (mod_obcs.F:569-580)
 *   LOG   bottom drag   (brough.F:88-98)
 *   **1.5 MY2.5 closure (vdif_q.F:260-263, :360-366)
 *
 * The same source is compiled for the host and for gfx950 with
 * -ffp-contract=off, and uses only IEEE add/sub/mul/div/sqrt, so host and
 * device results agree bit for bit.  Accuracy is a few ulp (checked against
 * libm in the test suite).
 *
 * Methods (standard, published):
 *   cos   Cody-Waite reduction by pi/2 with a three-part constant, then the
 *         Taylor series of sin/cos on [-pi/4, pi/4] to degree 19/18.
 *   expm1 reduction y = k ln2 + r, |r| <= ln2/2, Taylor series of e^r - 1 to
 *         degree 14, rescaling by 2^k.
 *   tanh  tanh(x) = -expm1(-2x) / (2 + expm1(-2x)) for x >= 0, odd symmetry.
 *   log   x = 2^e m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *         s = (m-1)/(m+1), atanh by its odd series to s^21; e ln2 added in
 *         two parts.
 *   x**1.5 = x * sqrt(x).
 */
#ifndef FV_DETMATH_H
#define FV_DETMATH_H

#if defined(__HIPCC__)
#define FV_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define FV_HD static inline
#endif

FV_HD double fv_ldexp_int(double x, int e) {
  /* exact scaling by 2^e for |e| <= 1000 using powers built by squaring-free loop */
  double f = 1.0;
  double b = e >= 0 ? 2.0 : 0.5;
  int n = e >= 0 ? e : -e;
  while (n) {
    if (n & 1) f = f * b;
    b = b * b;
    n >>= 1;
  }
  return x * f;
}

/* round-to-nearest integer of a double of moderate size */
FV_HD double fv_rint(double x) {
  const double big = 6755399441055744.0; /* 1.5 * 2^52 */
  return (x + big) - big;
}

FV_HD double fv_cos(double x) {
  /* pi/2 split in three parts: hi and mid carry 33 significant bits so that
   * n*hi and n*mid are exact for |n| < 2^20 */
  const double PIO2_HI = 1.57079632673412561417e+00;
  const double PIO2_MD = 6.07710050630396597660e-11;
  const double PIO2_LO = 2.02226624879595063154e-21;
  const double TWO_OVER_PI = 6.3661977236758138e-01;
  if (!(x - x == 0.0)) return x - x;  /* NaN or infinite argument: NaN */
  double n = fv_rint(x * TWO_OVER_PI);
  double r = ((x - n * PIO2_HI) - n * PIO2_MD) - n * PIO2_LO;
  long q = (long)n;
  int quad = (int)(q & 3);
  double r2 = r * r;
  /* cos(r) and sin(r), Taylor coefficients 1/k! */
  double c = 1.0 + r2 * (-5.0e-1 + r2 * (4.1666666666666664e-02 + r2 * (-1.3888888888888889e-03 +
             r2 * (2.4801587301587302e-05 + r2 * (-2.7557319223985893e-07 + r2 * (2.0876756987868100e-09 +
             r2 * (-1.1470745597729725e-11 + r2 * (4.7794773323873853e-14 + r2 * (-1.5619206968586225e-16)))))))));
  double s = r + r * r2 * (-1.6666666666666666e-01 + r2 * (8.3333333333333332e-03 + r2 * (-1.9841269841269841e-04 +
             r2 * (2.7557319223985893e-06 + r2 * (-2.5052108385441720e-08 + r2 * (1.6059043836821613e-10 +
             r2 * (-7.6471637318198164e-13 + r2 * (2.8114572543455206e-15 + r2 * (-8.2206352466243295e-18)))))))));
  switch (quad) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
  }
}

FV_HD double fv_expm1(double y) {
  if (y != y) return y;             /* NaN: no integer conversion of it below */
  if (y > 709.0) return 1.0 / 0.0;
  if (y < -40.0) return -1.0;
  const double LN2_HI = 6.93147180369123816490e-01; /* 33 bits */
  const double LN2_LO = 1.90821492927058770002e-10;
  const double INV_LN2 = 1.44269504088896338700e+00;
  double k = fv_rint(y * INV_LN2);
  double r = (y - k * LN2_HI) - k * LN2_LO;
  /* e^r - 1 = r + r^2/2! + ... + r^14/14! */
  double p = 1.0 / 87178291200.0;
  p = 1.0 / 6227020800.0 + r * p;
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
  double em1 = r + r * r * p;       /* e^r - 1 */
  int ki = (int)k;
  if (ki == 0) return em1;
  double two_k = fv_ldexp_int(1.0, ki);
  return two_k * em1 + (two_k - 1.0);
}

FV_HD double fv_tanh(double x) {
  double ax = x < 0 ? -x : x;
  double t;
  if (ax > 22.0) t = 1.0;
  else {
    double e = fv_expm1(-2.0 * ax);
    t = -e / (2.0 + e);
  }
  return x < 0 ? -t : t;
}

FV_HD double fv_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -1.0 / 0.0 : (x - x) / (x - x);
  if (x == 1.0 / 0.0) return x;
  /* decompose x = 2^e m with m in [sqrt(1/2), sqrt(2)) by exact halving/doubling */
  int e = 0;
  double m = x;
  while (m >= 1.4142135623730951) { m = m * 0.5; ++e; }
  while (m < 0.70710678118654757) { m = m * 2.0; --e; }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  /* atanh(s) = s + s^3/3 + ... + s^21/21 */
  double p = 1.0 / 21.0;
  p = 1.0 / 19.0 + s2 * p;
  p = 1.0 / 17.0 + s2 * p;
  p = 1.0 / 15.0 + s2 * p;
  p = 1.0 / 13.0 + s2 * p;
  p = 1.0 / 11.0 + s2 * p;
  p = 1.0 / 9.0 + s2 * p;
  p = 1.0 / 7.0 + s2 * p;
  p = 1.0 / 5.0 + s2 * p;
  p = 1.0 / 3.0 + s2 * p;
  double lm = 2.0 * s + 2.0 * s * (s2 * p);
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  double de = (double)e;
  return de * LN2_HI + (lm + de * LN2_LO);
}

FV_HD double fv_pow15(double x) { return x * sqrt(x); }

#endif /* FV_DETMATH_H */
