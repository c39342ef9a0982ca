/*
 * ORACLE — test infrastructure only. Nothing in the product links this file.
 *
 * Portable single-precision transcendentals used by the oracle when it runs in
 * "portable math" mode (the default). The reference calls glibc's sinf / cosf /
 * powf (ScalarCpuPathTracer.cpp:981-987, :1000-1019, :1037-1039, :1103); glibc's
 * results cannot be reproduced on a GPU. This file restates the same functions
 * from textbook methods so that a CPU and a GPU evaluating the same formulas
 * give bit-identical floats:
 *
 *   sin/cos : Cody & Waite (1980) three-part reduction by pi/2 in double, then
 *             the Taylor series of sin and cos on [-pi/4, pi/4] (terms through
 *             r^21 / r^20, truncation error < 1e-24), rounded once to float.
 *   pow     : x^y = exp(y * ln x).  ln via x = m * 2^e, m in [sqrt(1/2), sqrt(2)),
 *             ln m = 2 atanh(s), s = (m-1)/(m+1), odd series through s^29
 *             (truncation error < 1e-24); exp via Cody-Waite reduction by ln 2
 *             and the Taylor series through r^20 on |r| <= ln2/2, rounded once
 *             to float.
 *
 * All constants are 1/n!, 1/n and splits of pi/2 and ln 2, computed exactly in
 * rational arithmetic and rounded to double. Every expression is evaluated
 * unfused (the oracle is compiled with -ffp-contract=off), so results depend
 * only on IEEE-754 double add / mul / div.
 *
 * Pinned by tests/test_oracle_math.py against glibc over dense samplings of the
 * argument ranges the tracer uses (|error| <= 1 ulp everywhere, and equal to
 * glibc in all but a tiny fraction of cases), plus exact special cases.
 */
#ifndef PTO_MATH_H
#define PTO_MATH_H

#include <stdint.h>
#include <string.h>

static inline uint64_t pm_bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double pm_from_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* 2^k for k in [-1022, 1023]. */
static inline double pm_pow2i(int k) { return pm_from_bits((uint64_t)(k + 1023) << 52); }

/* Round half away from zero, |t| < 2^30. */
static inline int pm_rint(double t) { return (int)(t < 0.0 ? t - 0.5 : t + 0.5); }

/* sin(r) for |r| <= pi/4 (+ a little): r + r*r2*P(r2). */
static inline double pm_sin_kernel(double r) {
  const double r2 = r * r;
  double p = 0x1.71b8ef6dcf572p-66;   /* +1/21! */
  p = p * r2 - 0x1.2f49b46814157p-57; /* -1/19! */
  p = p * r2 + 0x1.952c77030ad4ap-49; /* +1/17! */
  p = p * r2 - 0x1.ae7f3e733b81fp-41; /* -1/15! */
  p = p * r2 + 0x1.6124613a86d09p-33; /* +1/13! */
  p = p * r2 - 0x1.ae64567f544e4p-26; /* -1/11! */
  p = p * r2 + 0x1.71de3a556c734p-19; /* +1/9!  */
  p = p * r2 - 0x1.a01a01a01a01ap-13; /* -1/7!  */
  p = p * r2 + 0x1.1111111111111p-7;  /* +1/5!  */
  p = p * r2 - 0x1.5555555555555p-3;  /* -1/3!  */
  return r + r * (r2 * p);
}

/* cos(r) for |r| <= pi/4 (+ a little): 1 + r2*Q(r2). */
static inline double pm_cos_kernel(double r) {
  const double r2 = r * r;
  double q = 0x1.e542ba4020225p-62;   /* +1/20! */
  q = q * r2 - 0x1.6827863b97d97p-53; /* -1/18! */
  q = q * r2 + 0x1.ae7f3e733b81fp-45; /* +1/16! */
  q = q * r2 - 0x1.93974a8c07c9dp-37; /* -1/14! */
  q = q * r2 + 0x1.1eed8eff8d898p-29; /* +1/12! */
  q = q * r2 - 0x1.27e4fb7789f5cp-22; /* -1/10! */
  q = q * r2 + 0x1.a01a01a01a01ap-16; /* +1/8!  */
  q = q * r2 - 0x1.6c16c16c16c17p-10; /* -1/6!  */
  q = q * r2 + 0x1.5555555555555p-5;  /* +1/4!  */
  q = q * r2 - 0x1.0000000000000p-1;  /* -1/2!  */
  return 1.0 + r2 * q;
}

/* Reduce x by pi/2: x = k*pi/2 + r. Valid for |x| < 2^22. */
static inline double pm_reduce_pio2(double x, int* quadrant) {
  const int k = pm_rint(x * 0x1.45f306dc9c883p-1); /* 2/pi */
  const double kd = (double)k;
  double r = x - kd * 0x1.921fb54000000p+0;         /* pi/2, leading 30 bits */
  r = r - kd * 0x1.10b4611800000p-30;               /* next 30 bits */
  r = r - kd * 0x1.313198a2e0370p-61;               /* tail */
  *quadrant = k & 3;
  return r;
}

static inline float pm_sinf(float xf) {
  const double x = (double)xf;
  if (!(x <= 4194304.0 && x >= -4194304.0)) {
    return (float)(x - x); /* NaN for inf/NaN; out-of-range args are not used by the tracer */
  }
  int q;
  const double r = pm_reduce_pio2(x, &q);
  double v;
  switch (q) {
    case 0: v = pm_sin_kernel(r); break;
    case 1: v = pm_cos_kernel(r); break;
    case 2: v = -pm_sin_kernel(r); break;
    default: v = -pm_cos_kernel(r); break;
  }
  return (float)v;
}

static inline float pm_cosf(float xf) {
  const double x = (double)xf;
  if (!(x <= 4194304.0 && x >= -4194304.0)) {
    return (float)(x - x);
  }
  int q;
  const double r = pm_reduce_pio2(x, &q);
  double v;
  switch (q) {
    case 0: v = pm_cos_kernel(r); break;
    case 1: v = -pm_sin_kernel(r); break;
    case 2: v = -pm_cos_kernel(r); break;
    default: v = pm_sin_kernel(r); break;
  }
  return (float)v;
}

/* Natural log of a positive, finite, normal double. */
static inline double pm_log(double x) {
  const uint64_t bits = pm_bits(x);
  int e = (int)((bits >> 52) & 0x7ffu) - 1023;
  double m = pm_from_bits((bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull); /* [1,2) */
  if (m > 0x1.6a09e667f3bcdp+0) { /* sqrt(2) */
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double l = 0x1.1a7b9611a7b96p-5;   /* 1/29 */
  l = l * s2 + 0x1.2f684bda12f68p-5; /* 1/27 */
  l = l * s2 + 0x1.47ae147ae147bp-5; /* 1/25 */
  l = l * s2 + 0x1.642c8590b2164p-5; /* 1/23 */
  l = l * s2 + 0x1.8618618618618p-5; /* 1/21 */
  l = l * s2 + 0x1.af286bca1af28p-5; /* 1/19 */
  l = l * s2 + 0x1.e1e1e1e1e1e1ep-5; /* 1/17 */
  l = l * s2 + 0x1.1111111111111p-4; /* 1/15 */
  l = l * s2 + 0x1.3b13b13b13b14p-4; /* 1/13 */
  l = l * s2 + 0x1.745d1745d1746p-4; /* 1/11 */
  l = l * s2 + 0x1.c71c71c71c71cp-4; /* 1/9  */
  l = l * s2 + 0x1.2492492492492p-3; /* 1/7  */
  l = l * s2 + 0x1.999999999999ap-3; /* 1/5  */
  l = l * s2 + 0x1.5555555555555p-2; /* 1/3  */
  const double two_s = s + s;
  const double ln_m = two_s + two_s * (s2 * l);
  const double ed = (double)e;
  return ed * 0x1.62e42fee00000p-1 + (ed * 0x1.a39ef35793c76p-33 + ln_m); /* ln2 hi + lo */
}

/* exp(z) for -110 <= z <= 100. */
static inline double pm_exp(double z) {
  const int k = pm_rint(z * 0x1.71547652b82fep+0); /* 1/ln2 */
  const double kd = (double)k;
  const double r = (z - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
  double p = 0x1.e542ba4020225p-62;   /* 1/20! */
  p = p * r + 0x1.2f49b46814157p-57;  /* 1/19! */
  p = p * r + 0x1.6827863b97d97p-53;  /* 1/18! */
  p = p * r + 0x1.952c77030ad4ap-49;  /* 1/17! */
  p = p * r + 0x1.ae7f3e733b81fp-45;  /* 1/16! */
  p = p * r + 0x1.ae7f3e733b81fp-41;  /* 1/15! */
  p = p * r + 0x1.93974a8c07c9dp-37;  /* 1/14! */
  p = p * r + 0x1.6124613a86d09p-33;  /* 1/13! */
  p = p * r + 0x1.1eed8eff8d898p-29;  /* 1/12! */
  p = p * r + 0x1.ae64567f544e4p-26;  /* 1/11! */
  p = p * r + 0x1.27e4fb7789f5cp-22;  /* 1/10! */
  p = p * r + 0x1.71de3a556c734p-19;  /* 1/9!  */
  p = p * r + 0x1.a01a01a01a01ap-16;  /* 1/8!  */
  p = p * r + 0x1.a01a01a01a01ap-13;  /* 1/7!  */
  p = p * r + 0x1.6c16c16c16c17p-10;  /* 1/6!  */
  p = p * r + 0x1.1111111111111p-7;   /* 1/5!  */
  p = p * r + 0x1.5555555555555p-5;   /* 1/4!  */
  p = p * r + 0x1.5555555555555p-3;   /* 1/3!  */
  p = p * r + 0x1.0000000000000p-1;   /* 1/2!  */
  const double e_r = 1.0 + (r + (r * r) * p);
  return e_r * pm_pow2i(k);
}

static inline float pm_powf(float xf, float yf) {
  if (xf != xf || yf != yf) {
    return xf + yf;
  }
  if (yf == 0.0f || xf == 1.0f) {
    return 1.0f;
  }
  const double x = (double)xf;
  const double y = (double)yf;
  if (xf < 0.0f) {
    return (float)((x - x) / (x - x)); /* NaN: negative bases are never used by the tracer */
  }
  if (xf == 0.0f) {
    return yf > 0.0f ? 0.0f : (float)(1.0 / x);
  }
  if (!(x <= 3.5e38)) { /* +inf */
    return yf > 0.0f ? xf : 0.0f;
  }
  const double z = y * pm_log(x);
  if (z > 100.0) {
    return (float)(z * 1e300); /* overflows to +inf in float */
  }
  if (z < -110.0) {
    return 0.0f;
  }
  return (float)pm_exp(z);
}

/* atan(x) for x >= 0 in double: x > 1 folds to pi/2 - atan(1/x); then atan(x) = atan(c) +
 * atan((x - c) / (1 + x c)) with c = k/16 the nearest sixteenth (|r| <= 1/32) and the odd Taylor
 * series of the remainder through r^15 (error < 1e-24). atan(k/16) computed to 50 digits. */
static const double pm_atan_k16[17] = {
    0x0.0p+0, 0x1.ff55bb72cfdeap-5, 0x1.fd5ba9aac2f6ep-4, 0x1.7b97b4bce5b02p-3, 0x1.f5b75f92c80ddp-3,
    0x1.362773707ebccp-2, 0x1.6f61941e4def1p-2, 0x1.a64eec3cc23fdp-2, 0x1.dac670561bb4fp-2,
    0x1.0657e94db30d0p-1, 0x1.1e00babdefeb4p-1, 0x1.345f01cce37bbp-1, 0x1.4978fa3269ee1p-1,
    0x1.5d58987169b18p-1, 0x1.700a7c5784634p-1, 0x1.819d0b7158a4dp-1, 0x1.921fb54442d18p-1};

static inline double pm_atan_pos(double x) {
  int inv = 0;
  if (x > 1.0) {
    x = 1.0 / x;
    inv = 1;
  }
  const int k = (int)(x * 16.0 + 0.5);
  const double c = (double)k * 0.0625;
  const double r = (x - c) / (1.0 + x * c);
  const double r2 = r * r;
  double p = -0x1.1111111111111p-4;  /* -1/15 */
  p = p * r2 + 0x1.3b13b13b13b14p-4; /* +1/13 */
  p = p * r2 - 0x1.745d1745d1746p-4; /* -1/11 */
  p = p * r2 + 0x1.c71c71c71c71cp-4; /* +1/9  */
  p = p * r2 - 0x1.2492492492492p-3; /* -1/7  */
  p = p * r2 + 0x1.999999999999ap-3; /* +1/5  */
  p = p * r2 - 0x1.5555555555555p-2; /* -1/3  */
  const double a = pm_atan_k16[k] + (r + r * (r2 * p));
  return inv ? (0x1.921fb54442d18p+0 - a) + 0x1.1a62633145c07p-54 : a; /* pi/2 hi + lo */
}

/* atan2 in double for finite arguments (C99 signs for zeros). */
static inline double pm_atan2(double y, double x) {
  const double ay = y < 0.0 ? -y : y, ax = x < 0.0 ? -x : x;
  double t;
  if (ax == 0.0 && ay == 0.0) {
    t = 0.0;
  } else if (ay <= ax) {
    t = pm_atan_pos(ay / ax);
  } else {
    t = (0x1.921fb54442d18p+0 - pm_atan_pos(ax / ay)) + 0x1.1a62633145c07p-54;
  }
  if (x < 0.0 || (x == 0.0 && 1.0 / x < 0.0)) t = (0x1.921fb54442d18p+1 - t) + 0x1.1a62633145c07p-53; /* pi - t */
  const int yneg = y < 0.0 || (y == 0.0 && 1.0 / y < 0.0);
  return yneg ? -t : t;
}

static inline float pm_atan2f(float y, float x) { return (float)pm_atan2((double)y, (double)x); }

/* sqrt of a non-negative double: correctly rounded float sqrt, then one Newton step in double. */
static inline double pm_sqrt_d(double v) {
  if (!(v > 0.0)) return 0.0;
  const double s0 = (double)sqrtf((float)v);
  if (!(s0 > 0.0)) return 0.0;
  return 0.5 * (s0 + v / s0);
}

/* acos(x) = atan2(sqrt((1 - x)(1 + x)), x) for x in [-1, 1]. */
static inline float pm_acosf(float xf) {
  const double x = (double)xf;
  return (float)pm_atan2(pm_sqrt_d((1.0 - x) * (1.0 + x)), x);
}

#endif /* PTO_MATH_H */
