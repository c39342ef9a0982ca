// pt_math.h — single-precision sin / cos / pow evaluated in IEEE double with a
// fixed operation order, so host code and gfx950 kernels produce bit-identical
// floats (built with -ffp-contract=off everywhere).
//
// The reference calls glibc sinf / cosf / powf on this path
// (src/pathtracer/ScalarCpuPathTracer.cpp:978-1020, :1034-1040, :1103). Their
// exact results are not reproducible on a GPU, so the path uses these instead:
//   sin/cos: Cody & Waite three-part reduction by pi/2, Taylor series on
//            [-pi/4, pi/4] through r^21 (sin) / r^20 (cos), one final rounding;
//   pow:     exp(y ln x); ln by m*2^e splitting and the atanh series in
//            s = (m-1)/(m+1) through s^29; exp by Cody & Waite reduction by ln 2
//            and the Taylor series through r^20; one final rounding.
// Coefficients are 1/n! and 1/n and splits of pi/2 and ln 2, rounded to double.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define PT_HD __host__ __device__
#else
#define PT_HD
#endif

namespace ptm {

PT_HD inline uint64_t dbits(double x) { return __builtin_bit_cast(uint64_t, x); }
PT_HD inline double dfrom(uint64_t u) { return __builtin_bit_cast(double, u); }
PT_HD inline double pow2i(int k) { return dfrom(static_cast<uint64_t>(k + 1023) << 52); }
PT_HD inline int rint_away(double t) { return static_cast<int>(t < 0.0 ? t - 0.5 : t + 0.5); }

PT_HD inline double sin_kernel(double r) {
  const double r2 = r * r;
  double p = 0x1.71b8ef6dcf572p-66;
  p = p * r2 - 0x1.2f49b46814157p-57;
  p = p * r2 + 0x1.952c77030ad4ap-49;
  p = p * r2 - 0x1.ae7f3e733b81fp-41;
  p = p * r2 + 0x1.6124613a86d09p-33;
  p = p * r2 - 0x1.ae64567f544e4p-26;
  p = p * r2 + 0x1.71de3a556c734p-19;
  p = p * r2 - 0x1.a01a01a01a01ap-13;
  p = p * r2 + 0x1.1111111111111p-7;
  p = p * r2 - 0x1.5555555555555p-3;
  return r + r * (r2 * p);
}

PT_HD inline double cos_kernel(double r) {
  const double r2 = r * r;
  double q = 0x1.e542ba4020225p-62;
  q = q * r2 - 0x1.6827863b97d97p-53;
  q = q * r2 + 0x1.ae7f3e733b81fp-45;
  q = q * r2 - 0x1.93974a8c07c9dp-37;
  q = q * r2 + 0x1.1eed8eff8d898p-29;
  q = q * r2 - 0x1.27e4fb7789f5cp-22;
  q = q * r2 + 0x1.a01a01a01a01ap-16;
  q = q * r2 - 0x1.6c16c16c16c17p-10;
  q = q * r2 + 0x1.5555555555555p-5;
  q = q * r2 - 0x1.0000000000000p-1;
  return 1.0 + r2 * q;
}

PT_HD inline double reduce_pio2(double x, int& quadrant) {
  const int k = rint_away(x * 0x1.45f306dc9c883p-1);
  const double kd = static_cast<double>(k);
  double r = x - kd * 0x1.921fb54000000p+0;
  r = r - kd * 0x1.10b4611800000p-30;
  r = r - kd * 0x1.313198a2e0370p-61;
  quadrant = k & 3;
  return r;
}

// sin and cos of the same float argument (one reduction).
PT_HD inline void sincosf_p(float xf, float& s_out, float& c_out) {
  const double x = static_cast<double>(xf);
  if (!(x <= 4194304.0 && x >= -4194304.0)) {
    s_out = static_cast<float>(x - x);
    c_out = s_out;
    return;
  }
  int q;
  const double r = reduce_pio2(x, q);
  const double s = sin_kernel(r);
  const double c = cos_kernel(r);
  double sv, cv;
  switch (q) {
    case 0: sv = s; cv = c; break;
    case 1: sv = c; cv = -s; break;
    case 2: sv = -s; cv = -c; break;
    default: sv = -c; cv = s; break;
  }
  s_out = static_cast<float>(sv);
  c_out = static_cast<float>(cv);
}

PT_HD inline float sinf_p(float x) { float s, c; sincosf_p(x, s, c); return s; }
PT_HD inline float cosf_p(float x) { float s, c; sincosf_p(x, s, c); return c; }

PT_HD inline double log_p(double x) {
  const uint64_t bits = dbits(x);
  int e = static_cast<int>((bits >> 52) & 0x7ffu) - 1023;
  double m = dfrom((bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 0x1.6a09e667f3bcdp+0) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double l = 0x1.1a7b9611a7b96p-5;
  l = l * s2 + 0x1.2f684bda12f68p-5;
  l = l * s2 + 0x1.47ae147ae147bp-5;
  l = l * s2 + 0x1.642c8590b2164p-5;
  l = l * s2 + 0x1.8618618618618p-5;
  l = l * s2 + 0x1.af286bca1af28p-5;
  l = l * s2 + 0x1.e1e1e1e1e1e1ep-5;
  l = l * s2 + 0x1.1111111111111p-4;
  l = l * s2 + 0x1.3b13b13b13b14p-4;
  l = l * s2 + 0x1.745d1745d1746p-4;
  l = l * s2 + 0x1.c71c71c71c71cp-4;
  l = l * s2 + 0x1.2492492492492p-3;
  l = l * s2 + 0x1.999999999999ap-3;
  l = l * s2 + 0x1.5555555555555p-2;
  const double two_s = s + s;
  const double ln_m = two_s + two_s * (s2 * l);
  const double ed = static_cast<double>(e);
  return ed * 0x1.62e42fee00000p-1 + (ed * 0x1.a39ef35793c76p-33 + ln_m);
}

PT_HD inline double exp_p(double z) {
  const int k = rint_away(z * 0x1.71547652b82fep+0);
  const double kd = static_cast<double>(k);
  const double r = (z - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
  double p = 0x1.e542ba4020225p-62;
  p = p * r + 0x1.2f49b46814157p-57;
  p = p * r + 0x1.6827863b97d97p-53;
  p = p * r + 0x1.952c77030ad4ap-49;
  p = p * r + 0x1.ae7f3e733b81fp-45;
  p = p * r + 0x1.ae7f3e733b81fp-41;
  p = p * r + 0x1.93974a8c07c9dp-37;
  p = p * r + 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0x1.0000000000000p-1;
  const double e_r = 1.0 + (r + (r * r) * p);
  return e_r * pow2i(k);
}

PT_HD inline float powf_p(float xf, float yf) {
  if (xf != xf || yf != yf) return xf + yf;
  if (yf == 0.0f || xf == 1.0f) return 1.0f;
  const double x = static_cast<double>(xf);
  const double y = static_cast<double>(yf);
  if (xf < 0.0f) return static_cast<float>((x - x) / (x - x));
  if (xf == 0.0f) return yf > 0.0f ? 0.0f : static_cast<float>(1.0 / x);
  if (!(x <= 3.5e38)) return yf > 0.0f ? xf : 0.0f;
  const double z = y * log_p(x);
  if (z > 100.0) return static_cast<float>(z * 1e300);
  if (z < -110.0) return 0.0f;
  return static_cast<float>(exp_p(z));
}

// atan(x), x >= 0: fold x > 1 to pi/2 - atan(1/x), then atan(x) = atan(k/16) + atan((x - c)/(1 + x c))
// with the nearest sixteenth c (|r| <= 1/32) and the odd Taylor series through r^15.
PT_HD inline double atan_pos(double x) {
  constexpr double kAtanK16[17] = {
      0x0.0p+0, 0x1.ff55bb72cfdeap-5, 0x1.fd5ba9aac2f6ep-4, 0x1.7b97b4bce5b02p-3, 0x1.f5b75f92c80ddp-3,
      0x1.362773707ebccp-2, 0x1.6f61941e4def1p-2, 0x1.a64eec3cc23fdp-2, 0x1.dac670561bb4fp-2,
      0x1.0657e94db30d0p-1, 0x1.1e00babdefeb4p-1, 0x1.345f01cce37bbp-1, 0x1.4978fa3269ee1p-1,
      0x1.5d58987169b18p-1, 0x1.700a7c5784634p-1, 0x1.819d0b7158a4dp-1, 0x1.921fb54442d18p-1};
  bool inv = false;
  if (x > 1.0) {
    x = 1.0 / x;
    inv = true;
  }
  const int k = static_cast<int>(x * 16.0 + 0.5);
  const double c = static_cast<double>(k) * 0.0625;
  const double r = (x - c) / (1.0 + x * c);
  const double r2 = r * r;
  double p = -0x1.1111111111111p-4;
  p = p * r2 + 0x1.3b13b13b13b14p-4;
  p = p * r2 - 0x1.745d1745d1746p-4;
  p = p * r2 + 0x1.c71c71c71c71cp-4;
  p = p * r2 - 0x1.2492492492492p-3;
  p = p * r2 + 0x1.999999999999ap-3;
  p = p * r2 - 0x1.5555555555555p-2;
  const double a = kAtanK16[k] + (r + r * (r2 * p));
  return inv ? (0x1.921fb54442d18p+0 - a) + 0x1.1a62633145c07p-54 : a;
}

PT_HD inline double atan2_d(double y, double x) {
  const double ay = y < 0.0 ? -y : y, ax = x < 0.0 ? -x : x;
  double t;
  if (ax == 0.0 && ay == 0.0) t = 0.0;
  else if (ay <= ax) t = atan_pos(ay / ax);
  else t = (0x1.921fb54442d18p+0 - atan_pos(ax / ay)) + 0x1.1a62633145c07p-54;
  if (x < 0.0 || (x == 0.0 && 1.0 / x < 0.0)) t = (0x1.921fb54442d18p+1 - t) + 0x1.1a62633145c07p-53;
  const bool yneg = y < 0.0 || (y == 0.0 && 1.0 / y < 0.0);
  return yneg ? -t : t;
}

PT_HD inline float atan2f_p(float y, float x) {
  return static_cast<float>(atan2_d(static_cast<double>(y), static_cast<double>(x)));
}

// sqrt of a non-negative double: correctly rounded fp32 sqrt, then one Newton step in double.
PT_HD inline double sqrt_d(double v) {
  if (!(v > 0.0)) return 0.0;
  const double s0 = static_cast<double>(__builtin_sqrtf(static_cast<float>(v)));
  if (!(s0 > 0.0)) return 0.0;
  return 0.5 * (s0 + v / s0);
}

PT_HD inline float acosf_p(float xf) {
  const double x = static_cast<double>(xf);
  return static_cast<float>(atan2_d(sqrt_d((1.0 - x) * (1.0 + x)), x));
}

}  // namespace ptm
