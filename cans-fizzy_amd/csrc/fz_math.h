// FP64 exp, expm1, log, tanh, sin and cos for the hot-path kernels.
//
// Why not the device math library: the reference's phase field goes through
// phi = eps*log((psi+eps)/(1-psi+eps)) (src/acdi.f90:244), which maps a
// one-ulp difference of psi in the saturated bulk (psi within 1e-14 of 0 or 1)
// to an O(eps) change of phi, hence of the Youngs normals and curvature and
// of the surface-tension force one step later. Two builds whose tanh/log
// differ by one ulp therefore drift apart at O(1e-1) within a few steps. These
// routines are written out with +,-,*,/ and fused multiply-add (Cody-Waite
// reduction, Taylor/atanh series in Horner form, within about two ulps);
// every operation is correctly rounded, so every build that uses them
// rounds identically. The CPU oracle carries its own copy of the same
// method, and tests pin both against the C library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define FZ_HD __host__ __device__ inline

// One Horner step fma(r, p, c) with the coefficient c as a scalar operand.
// VOP3 encodings take no 64-bit literal, and the compiler otherwise copies
// each coefficient into two vector registers to use the two-address v_fmac
// (two v_mov per step, a fifth of the phase-field kernels' vector
// instructions). The same fused operation, so the same result.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline double fz_hstep(double r, double p, double c) {
  double o;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(o) : "v"(r), "v"(p), "s"(c));
  return o;
}
#else
inline double fz_hstep(double r, double p, double c) { return __builtin_fma(r, p, c); }
#endif

namespace fzm {

FZ_HD double bits_to_double(uint64_t b) { return __builtin_bit_cast(double, b); }
FZ_HD uint64_t double_to_bits(double x) { return __builtin_bit_cast(uint64_t, x); }

// x * 2^k for integer k in [-2044, 2046] (two exact power-of-two factors)
FZ_HD double scale2(double x, int k) {
  if (k > 1023) {
    x *= bits_to_double((uint64_t)(1023 + 1023) << 52);
    k -= 1023;
  } else if (k < -1022) {
    x *= bits_to_double((uint64_t)(1023 - 1022) << 52);
    k += 1022;
  }
  return x * bits_to_double((uint64_t)(k + 1023) << 52);
}

constexpr double LN2_HI = 0.6931467056274414;      // ln 2 with 32 trailing zero bits
constexpr double LN2_LO = 4.7493250390316726e-07;  // ln 2 - LN2_HI
constexpr double INV_LN2 = 1.4426950408889634;

constexpr double LN2_HALF = 0.34657359027997264;  // ln 2 / 2

// exp(r) - 1 for |r| <= ln2/2: Taylor series to r^14, Horner with fma
FZ_HD double expm1_kernel(double r) {
  double p = 1.1470745597729725e-11;
  p = fz_hstep(r, p, 1.6059043836821613e-10);
  p = fz_hstep(r, p, 2.08767569878681e-09);
  p = fz_hstep(r, p, 2.505210838544172e-08);
  p = fz_hstep(r, p, 2.755731922398589e-07);
  p = fz_hstep(r, p, 2.7557319223985893e-06);
  p = fz_hstep(r, p, 2.48015873015873e-05);
  p = fz_hstep(r, p, 0.0001984126984126984);
  p = fz_hstep(r, p, 0.001388888888888889);
  p = fz_hstep(r, p, 0.008333333333333333);
  p = fz_hstep(r, p, 0.041666666666666664);
  p = fz_hstep(r, p, 0.16666666666666666);
  p = __builtin_fma(r, p, 0.5);
  return __builtin_fma(r * r, p, r);
}

// exp(r) - 1 for |r| <= 1.05: Taylor series to r^20, Horner with fma (tanh's small arguments)
FZ_HD double expm1_kernel20(double r) {
  double p = 4.110317623312165e-19;
  p = fz_hstep(r, p, 8.22063524662433e-18);
  p = fz_hstep(r, p, 1.5619206968586225e-16);
  p = fz_hstep(r, p, 2.8114572543455206e-15);
  p = fz_hstep(r, p, 4.779477332387385e-14);
  p = fz_hstep(r, p, 7.647163731819816e-13);
  p = fz_hstep(r, p, 1.1470745597729725e-11);
  p = fz_hstep(r, p, 1.6059043836821613e-10);
  p = fz_hstep(r, p, 2.08767569878681e-09);
  p = fz_hstep(r, p, 2.505210838544172e-08);
  p = fz_hstep(r, p, 2.755731922398589e-07);
  p = fz_hstep(r, p, 2.7557319223985893e-06);
  p = fz_hstep(r, p, 2.48015873015873e-05);
  p = fz_hstep(r, p, 0.0001984126984126984);
  p = fz_hstep(r, p, 0.001388888888888889);
  p = fz_hstep(r, p, 0.008333333333333333);
  p = fz_hstep(r, p, 0.041666666666666664);
  p = fz_hstep(r, p, 0.16666666666666666);
  p = __builtin_fma(r, p, 0.5);
  return __builtin_fma(r * r, p, r);
}

// x = k ln2 + r, |r| <= ln2/2 (k*LN2_HI is exact)
FZ_HD double reduce_ln2(double x, int *k) {
  const double kf = __builtin_floor(x * INV_LN2 + 0.5);
  *k = (int)kf;
  return __builtin_fma(-kf, LN2_LO, x - kf * LN2_HI);
}

FZ_HD double exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_huge_val();
  if (x < -745.1332191019412) return 0.0;
  int k;
  const double r = reduce_ln2(x, &k);
  return scale2(1.0 + expm1_kernel(r), k);
}

// expm1(x) = 2^k (1 + expm1(r)) - 1 = 2^k expm1(r) + (2^k - 1); both terms
// are exact scalings or exact differences, so one rounding remains
FZ_HD double expm1(double x) {
  if (x != x) return x;
  const double a = x < 0.0 ? -x : x;
  if (a <= LN2_HALF) return expm1_kernel(x);
  if (x > 709.782712893384) return __builtin_huge_val();
  if (x < -40.0) return -1.0;
  int k;
  const double e = expm1_kernel(reduce_ln2(x, &k));
  return scale2(e, k) + (scale2(1.0, k) - 1.0);
}

// log(x): x = 2^k m, m in [sqrt(2)/2, sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1)
FZ_HD double log(double x) {
  if (x != x) return x;
  if (x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_huge_val();
  if (x == __builtin_huge_val()) return x;
  int k = 0;
  if (x < 2.2250738585072014e-308) {  // subnormal
    x *= 18014398509481984.0;          // 2^54
    k = -54;
  }
  const uint64_t b = double_to_bits(x);
  k += (int)((b >> 52) & 0x7ff) - 1023;
  double m = bits_to_double((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1,2)
  if (m > 1.4142135623730951) {
    m = 0.5 * m;
    k += 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 0.08695652173913043;
  p = fz_hstep(z, p, 0.09523809523809523);
  p = fz_hstep(z, p, 0.10526315789473684);
  p = fz_hstep(z, p, 0.11764705882352941);
  p = fz_hstep(z, p, 0.13333333333333333);
  p = fz_hstep(z, p, 0.15384615384615385);
  p = fz_hstep(z, p, 0.18181818181818182);
  p = fz_hstep(z, p, 0.2222222222222222);
  p = fz_hstep(z, p, 0.2857142857142857);
  p = fz_hstep(z, p, 0.4);
  p = fz_hstep(z, p, 0.6666666666666666);
  const double lm = __builtin_fma(s * z, p, 2.0 * s);
  const double kd = (double)k;
  return __builtin_fma(kd, LN2_HI, __builtin_fma(kd, LN2_LO, lm));
}

// n / d by the device's own correctly rounded FP64 division sequence (the
// reciprocal, two Newton refinements, the quotient and one correction)
// without its range scaling and special-value fix-up, which leave the
// quotient unchanged when n and d lie in [2^-300, 2^300] in magnitude: there
// the result is n / d to the last bit. Outside that range it is meaningless.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline double div_inrange(double n, double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-d, y, 1.0);
  y = __builtin_fma(y, e, y);
  const double q = n * y;
  const double rm = __builtin_fma(-d, q, n);
  return __builtin_fma(rm, y, q);
}
#else
inline double div_inrange(double n, double d) { return n / d; }
#endif

// tanh(x) = expm1(2a) / (expm1(2a) + 2), a = |x|, without branches, so that
// the lanes of a wavefront stay on one instruction stream and the compiler
// can interleave several evaluations: expm1(2a) = 2^k expm1(r) + (2^k - 1)
// with 2a = k ln2 + r, |r| <= ln2/2 (k <= 64 after clamping a at 22, so both
// terms are exact); tanh = 1 above 22 and a itself below 2^-28
FZ_HD double tanh(double x) {
  const double a = x < 0.0 ? -x : x;
  const double y = 2.0 * (a < 22.0 ? a : 22.0);
  const double kf = __builtin_floor(y * INV_LN2 + 0.5);
  const double r = __builtin_fma(-kf, LN2_LO, y - kf * LN2_HI);
  const double e = expm1_kernel(r);
  const double p2 = bits_to_double((uint64_t)((int)kf + 1023) << 52);  // 2^k
  const double em1 = e * p2 + (p2 - 1.0);
  const double t = div_inrange(em1, em1 + 2.0);  // both in [7e-9, 2^64] where t is used
  double res = a > 22.0 ? 1.0 : t;
  res = a < 3.725290298461914e-09 ? a : res;
  res = x < 0.0 ? -res : res;
  return x != x ? x : res;
}

constexpr double PIO2_1 = 1.5707963109016418;     // pi/2 in three 26-bit pieces
constexpr double PIO2_2 = 1.5893254712295857e-08;
constexpr double PIO2_3 = 6.123233995736766e-17;
constexpr double TWO_OVER_PI = 0.6366197723675814;

FZ_HD double sin_kernel(double r) {  // |r| <= pi/4
  const double z = r * r;
  double p = -7.647163731819816e-13;
  p = fz_hstep(z, p, 1.6059043836821613e-10);
  p = fz_hstep(z, p, -2.505210838544172e-08);
  p = fz_hstep(z, p, 2.7557319223985893e-06);
  p = fz_hstep(z, p, -0.0001984126984126984);
  p = fz_hstep(z, p, 0.008333333333333333);
  p = fz_hstep(z, p, -0.16666666666666666);
  return __builtin_fma(r * z, p, r);
}
FZ_HD double cos_kernel(double r) {
  const double z = r * r;
  double p = 4.779477332387385e-14;
  p = fz_hstep(z, p, -1.1470745597729725e-11);
  p = fz_hstep(z, p, 2.08767569878681e-09);
  p = fz_hstep(z, p, -2.755731922398589e-07);
  p = fz_hstep(z, p, 2.48015873015873e-05);
  p = fz_hstep(z, p, -0.001388888888888889);
  p = fz_hstep(z, p, 0.041666666666666664);
  return __builtin_fma(z * z, p, 1.0 - 0.5 * z);
}
// quadrant reduction for |x| < 2^20 (kf * PIO2_1 is exact)
FZ_HD double reduce_pio2(double x, int *q) {
  const double kf = __builtin_floor(x * TWO_OVER_PI + 0.5);
  *q = ((int)kf) & 3;
  return __builtin_fma(-kf, PIO2_3, __builtin_fma(-kf, PIO2_2, x - kf * PIO2_1));
}
FZ_HD double sin(double x) {
  int q;
  const double r = reduce_pio2(x, &q);
  switch (q) {
    case 0: return sin_kernel(r);
    case 1: return cos_kernel(r);
    case 2: return -sin_kernel(r);
    default: return -cos_kernel(r);
  }
}
FZ_HD double cos(double x) {
  int q;
  const double r = reduce_pio2(x, &q);
  switch (q) {
    case 0: return cos_kernel(r);
    case 1: return -sin_kernel(r);
    case 2: return -cos_kernel(r);
    default: return sin_kernel(r);
  }
}

}  // namespace fzm
