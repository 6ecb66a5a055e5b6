/*
 * TEST INFRASTRUCTURE ONLY (part of the CPU oracle).
 *
 * exp, expm1, log, tanh, sin, cos in FP64 with +,-,*,/ and fma only:
 * Cody-Waite argument reduction and Taylor/atanh series in Horner form
 * (within about two ulps). The oracle uses these instead of the C library
 * so that a build of the same method on the GPU rounds identically (every
 * operation, fma included, is correctly rounded): the reference's
 * phi = eps*log((psi+eps)/(1-psi+eps)) (src/acdi.f90:244) turns one-ulp
 * differences of psi in the saturated bulk into O(eps) changes of phi, so
 * two math libraries that differ by one ulp make otherwise identical runs
 * drift apart within a few steps. tests/test_oracle_math.py pins these
 * functions against numpy/libm in extended precision.
 */
#ifndef FZ_ORACLE_MATH_H
#define FZ_ORACLE_MATH_H
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline double om_bits_to_double(uint64_t b) { double x; memcpy(&x, &b, 8); return x; }
static inline uint64_t om_double_to_bits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }

static inline double om_scale2(double x, int k) {
  if (k > 1023) { x *= om_bits_to_double((uint64_t)(1023 + 1023) << 52); k -= 1023; }
  else if (k < -1022) { x *= om_bits_to_double((uint64_t)(1023 - 1022) << 52); k += 1022; }
  return x * om_bits_to_double((uint64_t)(k + 1023) << 52);
}

#define OM_LN2_HI 0.6931467056274414
#define OM_LN2_LO 4.7493250390316726e-07
#define OM_INV_LN2 1.4426950408889634

#define OM_LN2_HALF 0.34657359027997264 /* ln 2 / 2 */

/* exp(r) - 1 for |r| <= ln2/2: Taylor series to r^14, Horner with fma */
static inline double om_expm1_kernel(double r) {
  double p = 1.1470745597729725e-11;
  p = fma(r, p, 1.6059043836821613e-10);
  p = fma(r, p, 2.08767569878681e-09);
  p = fma(r, p, 2.505210838544172e-08);
  p = fma(r, p, 2.755731922398589e-07);
  p = fma(r, p, 2.7557319223985893e-06);
  p = fma(r, p, 2.48015873015873e-05);
  p = fma(r, p, 0.0001984126984126984);
  p = fma(r, p, 0.001388888888888889);
  p = fma(r, p, 0.008333333333333333);
  p = fma(r, p, 0.041666666666666664);
  p = fma(r, p, 0.16666666666666666);
  p = fma(r, p, 0.5);
  return fma(r * r, p, r);
}

/* exp(r) - 1 for |r| <= 1.05: Taylor series to r^20, Horner with fma (tanh's small arguments) */
static inline double om_expm1_kernel20(double r) {
  double p = 4.110317623312165e-19;
  p = fma(r, p, 8.22063524662433e-18);
  p = fma(r, p, 1.5619206968586225e-16);
  p = fma(r, p, 2.8114572543455206e-15);
  p = fma(r, p, 4.779477332387385e-14);
  p = fma(r, p, 7.647163731819816e-13);
  p = fma(r, p, 1.1470745597729725e-11);
  p = fma(r, p, 1.6059043836821613e-10);
  p = fma(r, p, 2.08767569878681e-09);
  p = fma(r, p, 2.505210838544172e-08);
  p = fma(r, p, 2.755731922398589e-07);
  p = fma(r, p, 2.7557319223985893e-06);
  p = fma(r, p, 2.48015873015873e-05);
  p = fma(r, p, 0.0001984126984126984);
  p = fma(r, p, 0.001388888888888889);
  p = fma(r, p, 0.008333333333333333);
  p = fma(r, p, 0.041666666666666664);
  p = fma(r, p, 0.16666666666666666);
  p = fma(r, p, 0.5);
  return fma(r * r, p, r);
}

/* x = k ln2 + r, |r| <= ln2/2 (k*OM_LN2_HI is exact) */
static inline double om_reduce_ln2(double x, int *k) {
  const double kf = floor(x * OM_INV_LN2 + 0.5);
  *k = (int)kf;
  return fma(-kf, OM_LN2_LO, x - kf * OM_LN2_HI);
}

static inline double om_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.1332191019412) return 0.0;
  int k;
  const double r = om_reduce_ln2(x, &k);
  return om_scale2(1.0 + om_expm1_kernel(r), k);
}

/* expm1(x) = 2^k (1 + expm1(r)) - 1 = 2^k expm1(r) + (2^k - 1); both terms */
/* are exact scalings or exact differences, so one rounding remains */
static inline double om_expm1(double x) {
  if (x != x) return x;
  const double a = x < 0.0 ? -x : x;
  if (a <= OM_LN2_HALF) return om_expm1_kernel(x);
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -40.0) return -1.0;
  int k;
  const double e = om_expm1_kernel(om_reduce_ln2(x, &k));
  return om_scale2(e, k) + (om_scale2(1.0, k) - 1.0);
}

/* log(x): x = 2^k m, m in [sqrt(2)/2, sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1) */
static inline double om_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int k = 0;
  if (x < 2.2250738585072014e-308) {  /* subnormal */
    x *= 18014398509481984.0;          /* 2^54 */
    k = -54;
  }
  const uint64_t b = om_double_to_bits(x);
  k += (int)((b >> 52) & 0x7ff) - 1023;
  double m = om_bits_to_double((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  /* [1,2) */
  if (m > 1.4142135623730951) {
    m = 0.5 * m;
    k += 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 0.08695652173913043;
  p = fma(z, p, 0.09523809523809523);
  p = fma(z, p, 0.10526315789473684);
  p = fma(z, p, 0.11764705882352941);
  p = fma(z, p, 0.13333333333333333);
  p = fma(z, p, 0.15384615384615385);
  p = fma(z, p, 0.18181818181818182);
  p = fma(z, p, 0.2222222222222222);
  p = fma(z, p, 0.2857142857142857);
  p = fma(z, p, 0.4);
  p = fma(z, p, 0.6666666666666666);
  const double lm = fma(s * z, p, 2.0 * s);
  const double kd = (double)k;
  return fma(kd, OM_LN2_HI, fma(kd, OM_LN2_LO, lm));
}

/* tanh(x) = expm1(2a) / (expm1(2a) + 2), a = |x|, evaluated for every
 * argument the same way (no branches in the device build): expm1(2a) =
 * 2^k expm1(r) + (2^k - 1) with 2a = k ln2 + r, |r| <= ln2/2 (k <= 64 after
 * clamping a at 22, so both terms are exact); 1 above 22, a below 2^-28 */
static inline double om_tanh(double x) {
  const double a = x < 0.0 ? -x : x;
  const double y = 2.0 * (a < 22.0 ? a : 22.0);
  const double kf = floor(y * OM_INV_LN2 + 0.5);
  const double r = fma(-kf, OM_LN2_LO, y - kf * OM_LN2_HI);
  const double e = om_expm1_kernel(r);
  const double p2 = om_bits_to_double((uint64_t)((int)kf + 1023) << 52); /* 2^k */
  const double em1 = e * p2 + (p2 - 1.0);
  const double t = em1 / (em1 + 2.0);
  double res = a > 22.0 ? 1.0 : t;
  res = a < 3.725290298461914e-09 ? a : res;
  res = x < 0.0 ? -res : res;
  return x != x ? x : res;
}

#define OM_PIO2_1 1.5707963109016418 /* pi/2 in three 26-bit pieces */
#define OM_PIO2_2 1.5893254712295857e-08
#define OM_PIO2_3 6.123233995736766e-17
#define OM_TWO_OVER_PI 0.6366197723675814

static inline double om_sin_kernel(double r) {  /* |r| <= pi/4 */
  const double z = r * r;
  double p = -7.647163731819816e-13;
  p = fma(z, p, 1.6059043836821613e-10);
  p = fma(z, p, -2.505210838544172e-08);
  p = fma(z, p, 2.7557319223985893e-06);
  p = fma(z, p, -0.0001984126984126984);
  p = fma(z, p, 0.008333333333333333);
  p = fma(z, p, -0.16666666666666666);
  return fma(r * z, p, r);
}
static inline double om_cos_kernel(double r) {
  const double z = r * r;
  double p = 4.779477332387385e-14;
  p = fma(z, p, -1.1470745597729725e-11);
  p = fma(z, p, 2.08767569878681e-09);
  p = fma(z, p, -2.755731922398589e-07);
  p = fma(z, p, 2.48015873015873e-05);
  p = fma(z, p, -0.001388888888888889);
  p = fma(z, p, 0.041666666666666664);
  return fma(z * z, p, 1.0 - 0.5 * z);
}
/* quadrant reduction for |x| < 2^20 (kf * OM_PIO2_1 is exact) */
static inline double om_reduce_pio2(double x, int *q) {
  const double kf = floor(x * OM_TWO_OVER_PI + 0.5);
  *q = ((int)kf) & 3;
  return fma(-kf, OM_PIO2_3, fma(-kf, OM_PIO2_2, x - kf * OM_PIO2_1));
}
static inline double om_sin(double x) {
  int q;
  const double r = om_reduce_pio2(x, &q);
  switch (q) {
  case 0: return om_sin_kernel(r);
  case 1: return om_cos_kernel(r);
  case 2: return -om_sin_kernel(r);
  default: return -om_cos_kernel(r);
  }
}
static inline double om_cos(double x) {
  int q;
  const double r = om_reduce_pio2(x, &q);
  switch (q) {
  case 0: return om_cos_kernel(r);
  case 1: return -om_sin_kernel(r);
  case 2: return -om_cos_kernel(r);
  default: return om_sin_kernel(r);
  }
}
#endif
