// Deterministic elementary functions for the renderer, identical bit-for-bit on
// the host (gcc / clang, x86-64 SSE) and on gfx950.
//
// The reference calls std::sin/cos/tan/atan/atan2/acos/pow/fmod on float
// arguments (core/src/Sample.cpp:5-42, 77-117; core/src/BRDF.cpp:105-109,
// 175-179; core/include/Hasty/VMath.h:682-699; core/include/Hasty/Environment.h:
// 117-166; core/src/Camera.cpp:53-76, 83-98; core/include/Hasty/PathTracing.h:
// 46-59).  glibc and the device math library differ by an ulp now and then, and
// one flipped comparison (a rejection test, a lobe choice) changes a whole path.
// So every transcendental used on the path goes through this header instead:
// each float function evaluates in double precision using only +, -, *, /,
// floor and correctly rounded sqrt, then rounds once to float.  With FP
// contraction disabled (-ffp-contract=off on both compilers) those operations
// are IEEE-754 correctly rounded everywhere, so the results are identical.
//
// Methods (all from the mathematics, coefficients are 1/n! or 1/n):
//   sin/cos : 3-term Cody-Waite reduction by pi/2, Taylor series on |r|<=pi/4.
//   atan    : reduction atan(x) = atan(c) + atan((x-c)/(1+xc)), c = k/8, then
//             the Gregory series on |z| <= 1/16.
//   log     : x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)).
//   exp     : x = k ln2 + r, Taylor series on |r| <= ln2/2, scale by 2^k.
//   fmod    : exact, via an exact double remainder.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define HM_HD __host__ __device__ inline
#else
#define HM_HD inline
#endif

namespace hm {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kInvPi = 1.0 / 3.14159265358979323846264338327950288;

HM_HD double dfloor(double x) { return __builtin_floor(x); }
HM_HD double dsqrt(double x) { return __builtin_sqrt(x); }
HM_HD float fsqrt(float x) { return __builtin_sqrtf(x); }
HM_HD bool is_nan(double x) { return x != x; }
HM_HD bool sign_bit(double x) { return __builtin_bit_cast(uint64_t, x) >> 63; }
HM_HD double make_pow2(int k) {  // 2^k for -1022 <= k <= 1023
  return __builtin_bit_cast(double, (uint64_t)(int64_t)(k + 1023) << 52);
}

// ---- sin / cos -------------------------------------------------------------
// pi/2 split into 33 + 33 + 53 significant bits so k*P1 and k*P2 are exact
// for |k| < 2^20.
constexpr double kPio2_1 = 0x1.921fb54400000p+0;
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;

// Series lengths: the float result needs about 40 correct bits (then it is
// the correctly rounded value except in ~2^-16 of cases, and within 1 ulp
// always); each series stops where its next term is below 2^-44 of the
// value over the reduced range.
HM_HD double sin_poly(double r) {  // r - r^3/3! + ... + r^13/13!  (|r| <= pi/4)
  double r2 = r * r;
  double p = 1.0 / 6227020800.0;
  p = p * r2 - 1.0 / 39916800.0;
  p = p * r2 + 1.0 / 362880.0;
  p = p * r2 - 1.0 / 5040.0;
  p = p * r2 + 1.0 / 120.0;
  p = p * r2 - 1.0 / 6.0;
  return r + r * (r2 * p);
}
HM_HD double cos_poly(double r) {  // 1 - r^2/2! + ... - r^14/14!  (|r| <= pi/4)
  double r2 = r * r;
  double p = -1.0 / 87178291200.0;
  p = p * r2 + 1.0 / 479001600.0;
  p = p * r2 - 1.0 / 3628800.0;
  p = p * r2 + 1.0 / 40320.0;
  p = p * r2 - 1.0 / 720.0;
  p = p * r2 + 1.0 / 24.0;
  p = p * r2 - 0.5;
  return 1.0 + r2 * p;
}
// Reduces x to r in [-pi/4, pi/4] and quadrant q = k mod 4.
HM_HD double reduce_pio2(double x, int* q) {
  double k = dfloor(x * (2.0 / kPi) + 0.5);
  double r = x - k * kPio2_1;
  r = r - k * kPio2_2;
  r = r - k * kPio2_3;
  long long ki = (long long)k;
  *q = (int)(ki & 3);
  return r;
}
HM_HD double dsin(double x) {
  if (is_nan(x)) return x;
  int q;
  double r = reduce_pio2(x, &q);
  switch (q) {
    case 0: return sin_poly(r);
    case 1: return cos_poly(r);
    case 2: return -sin_poly(r);
    default: return -cos_poly(r);
  }
}
HM_HD double dcos(double x) {
  if (is_nan(x)) return x;
  int q;
  double r = reduce_pio2(x, &q);
  switch (q) {
    case 0: return cos_poly(r);
    case 1: return -sin_poly(r);
    case 2: return -cos_poly(r);
    default: return sin_poly(r);
  }
}
HM_HD float sinf(float x) { return (float)dsin((double)x); }
HM_HD float cosf(float x) { return (float)dcos((double)x); }
// sinf(x) and cosf(x), bit for bit, from one reduction.
HM_HD void sincosf(float x, float* s, float* c) {
  if (x != x) {
    *s = *c = x;
    return;
  }
  int q;
  const double r = reduce_pio2((double)x, &q);
  const double sp = sin_poly(r), cp = cos_poly(r);
  switch (q) {
    case 0: *s = (float)sp, *c = (float)cp; break;
    case 1: *s = (float)cp, *c = (float)-sp; break;
    case 2: *s = (float)-sp, *c = (float)-cp; break;
    default: *s = (float)-cp, *c = (float)sp; break;
  }
}
HM_HD float tanf(float x) {
  double d = (double)x;
  return (float)(dsin(d) / dcos(d));
}

// ---- atan / atan2 / acos ---------------------------------------------------
// atan(k/8), k = 0..8, to double precision.
HM_HD double atan_table(int k) {
  switch (k) {
    case 0: return 0.0;
    case 1: return 0x1.fd5ba9aac2f6ep-4;
    case 2: return 0x1.f5b75f92c80ddp-3;
    case 3: return 0x1.6f61941e4def1p-2;
    case 4: return 0x1.dac670561bb4fp-2;
    case 5: return 0x1.1e00babdefeb4p-1;
    case 6: return 0x1.4978fa3269ee1p-1;
    case 7: return 0x1.700a7c5784634p-1;
    default: return 0x1.921fb54442d18p-1;
  }
}
// atan on [0, 1].
HM_HD double atan01(double a) {
  int k = (int)dfloor(a * 8.0 + 0.5);
  double c = k * 0.125;
  double z = (a - c) / (1.0 + a * c);
  double z2 = z * z;
  // z - z^3/3 + z^5/5 - ... - z^11/11 (|z| <= 1/16: next term < 2e-16 z)
  double p = -1.0 / 11.0;
  p = p * z2 + 1.0 / 9.0;
  p = p * z2 - 1.0 / 7.0;
  p = p * z2 + 1.0 / 5.0;
  p = p * z2 - 1.0 / 3.0;
  return atan_table(k) + (z + z * (z2 * p));
}
HM_HD double datan(double x) {
  if (is_nan(x)) return x;
  double a = x < 0.0 ? -x : x;
  double r;
  if (a <= 1.0) {
    r = atan01(a);
  } else if (a > 1e300) {
    r = kPi * 0.5;
  } else {
    r = kPi * 0.5 - atan01(1.0 / a);
  }
  return x < 0.0 ? -r : r;
}
// C99 atan2 semantics for finite arguments (signed zeros included).
HM_HD double datan2(double y, double x) {
  if (is_nan(x) || is_nan(y)) return x + y;
  if (x == 0.0 && y == 0.0) {
    double base = sign_bit(x) ? kPi : 0.0;
    return sign_bit(y) ? -base : base;
  }
  if (x == 0.0) return sign_bit(y) ? -kPi * 0.5 : kPi * 0.5;
  double t = datan(y / x);
  if (x > 0.0) return t;
  return sign_bit(y) ? t - kPi : t + kPi;
}
HM_HD float atanf(float x) { return (float)datan((double)x); }
HM_HD float atan2f(float y, float x) { return (float)datan2((double)y, (double)x); }
HM_HD float acosf(float x) {
  double d = (double)x;
  if (!(d >= -1.0 && d <= 1.0)) return __builtin_nanf("");
  double s = dsqrt((1.0 - d) * (1.0 + d));
  return (float)datan2(s, d);
}

// ---- log / exp / pow -------------------------------------------------------
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;

HM_HD double dlog(double x) {  // natural log, x > 0 finite
  if (is_nan(x) || x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_inf();
  if (x == __builtin_inf()) return x;
  uint64_t bits = __builtin_bit_cast(uint64_t, x);
  int e = (int)((bits >> 52) & 0x7ff);
  if (e == 0) {  // subnormal: scale up by 2^54
    x = x * 0x1p54;
    bits = __builtin_bit_cast(uint64_t, x);
    e = (int)((bits >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  double m = __builtin_bit_cast(double, (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  // 2*(s + s^3/3 + ... + s^17/17); |s| <= 0.1716 (next term < 3e-15 s)
  double p = 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  double lm = 2.0 * (s + s * (s2 * p));
  double de = (double)e;
  return (de * kLn2Hi + lm) + de * kLn2Lo;
}
HM_HD double dexp(double x) {
  if (is_nan(x)) return x;
  if (x > 709.0) return __builtin_inf();
  if (x < -745.2) return 0.0;
  double k = dfloor(x * 1.4426950408889634 + 0.5);
  double r = (x - k * kLn2Hi) - k * kLn2Lo;
  // e^r, |r| <= 0.35: Taylor to r^14/14! (next term < 5e-18)
  double p = 1.0 / 87178291200.0;
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  double er = 1.0 + r + r * (r * p);
  int ki = (int)k;
  if (ki < -1022) {  // gradual underflow in two steps
    return (er * make_pow2(ki + 600)) * make_pow2(-600);
  }
  if (ki > 1023) return (er * make_pow2(ki - 1)) * 2.0;
  return er * make_pow2(ki);
}
// powf for base > 0 (the renderer only calls it with a positive base).
HM_HD float powf_pos(float base, float e) {
  if (base == 1.0f) return 1.0f;
  if (e == 0.0f) return 1.0f;
  return (float)dexp((double)e * dlog((double)base));
}

// ---- fmod (exact) ----------------------------------------------------------
HM_HD float fmodf(float x, float y) {
  if (x != x || y != y || y == 0.0f || __builtin_isinf(x)) return __builtin_nanf("");
  if (__builtin_isinf(y)) return x;
  double ax = x < 0.0f ? -(double)x : (double)x;
  double ay = y < 0.0f ? -(double)y : (double)y;
  if (ax < ay) return x;
  // For float operands with ax/ay < 2^29 every step below is exact in double.
  double r;
  if (ax / ay < 536870912.0) {
    double n = dfloor(ax / ay);
    r = ax - n * ay;
    if (r < 0.0) r += ay;
    if (r >= ay) r -= ay;
  } else {  // long division on exact scaled copies of ay
    r = ax;
    while (r >= ay) {
      double s = ay;
      while (s * 2.0 <= r) s *= 2.0;
      r -= s;
    }
  }
  float fr = (float)r;
  return x < 0.0f ? -fr : fr;
}

}  // namespace hm
