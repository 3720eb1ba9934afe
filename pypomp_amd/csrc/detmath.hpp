// Deterministic elementary functions for the filter's state, weight and
// ancestry path.  Each is a fixed sequence of correctly rounded IEEE operations
// (+, -, *, /, sqrt, rint, ldexp) with contraction disabled, so any conforming
// host or device evaluates it to the same bits; a NumPy restatement of the same
// sequences serves as the oracle.  Accuracy is a few ulp (float) or ~1e-15
// relative (double), from these published constructions:
//   log:   argument reduction to m in [sqrt(1/2), sqrt(2)), log(m) = 2 atanh(s),
//          s = (m-1)/(m+1), truncated atanh series (Abramowitz & Stegun 4.1.27 /
//          fdlibm e_log.c structure);
//   log1p: Kahan's log(u) * x / (u - 1) with u = 1 + x (Goldberg, "What every
//          computer scientist should know about floating-point arithmetic", 1991);
//   exp:   Cody-Waite reduction by ln 2 with a split constant, Taylor polynomial,
//          exact scaling by 2^k (Cody & Waite 1980);
//   expm1: Kahan's (u - 1) * x / log(u) with u = exp(x) (Goldberg 1991);
//   tanh:  -t / (t + 2), t = expm1(-2|x|);
//   cos:   Cody-Waite reduction by pi/2 (three-part constant), Taylor sin/cos
//          polynomials by quadrant;
//   exp (double): the same construction in float64 with a degree-13 polynomial;
//   lgamma (double): upward shift to x >= 12 and Stirling's series with
//          Bernoulli terms to 1/x^9 (Abramowitz & Stegun 6.1.40).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MP_DM __host__ __device__ inline

namespace mp {

MP_DM float dm_bits_to_float(uint32_t b) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = b;
  return c.f;
}
MP_DM uint32_t dm_float_to_bits(float f) {
  union {
    uint32_t u;
    float f;
  } c;
  c.f = f;
  return c.u;
}

// 2^k as float for -126 <= k <= 127
MP_DM float dm_pow2(int k) { return dm_bits_to_float((uint32_t)(k + 127) << 23); }

// ---- float32 ----

// Correctly rounded float division in fewer instructions on the device, for
// the operand ranges stated at each use; the host keeps the IEEE division.
// Both were checked against the hardware's IEEE division on gfx950 over every
// input of those ranges.
//  div_core: the Newton-Raphson core of the division sequence without its
//    scaling and special-case steps (exact when a, b and a / b stay far from
//    overflow and underflow; used for the log's f / (2 + f), 2 + f in
//    [1.70, 2.42], all 2^23 cases exact).
//  div_by: Markstein's correction q + (a - b q) rb with rb = fl(1 / b) (exact
//    when a is 0 or a and a / b are finite and above 2^-100 in magnitude; the
//    remainder a - b q must not underflow).
MP_DM float div_core(float a, float b) {
#ifdef __HIP_DEVICE_COMPILE__
  float y = __builtin_amdgcn_rcpf(b);
  const float e = __builtin_fmaf(-b, y, 1.0f);
  y = __builtin_fmaf(e, y, y);
  float q = a * y;
  float r = __builtin_fmaf(-b, q, a);
  q = __builtin_fmaf(r, y, q);
  r = __builtin_fmaf(-b, q, a);
  return __builtin_fmaf(r, y, q);
#else
  return a / b;
#endif
}

MP_DM float div_by(float a, float b, float rb) {
#ifdef __HIP_DEVICE_COMPILE__
  const float q = a * rb;
  const float r = __builtin_fmaf(-b, q, a);
  return __builtin_fmaf(r, rb, q);
#else
  (void)rb;
  return a / b;
#endif
}

MP_DM float d_log(float x) {
#pragma clang fp contract(off)
  if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
  if (x == __builtin_inff()) return x;
  int e = 0;
  if (x < 1.17549435e-38f) {  // subnormal: scale into the normal range
    x = x * 8388608.0f;
    e = -23;
  }
  const uint32_t b = dm_float_to_bits(x);
  e += (int)((b >> 23) & 0xffu) - 127;
  float m = dm_bits_to_float((b & 0x7fffffu) | 0x3f800000u);  // [1, 2)
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  const float s = div_core(f, 2.0f + f);
  const float z = s * s;
  float p = 0.111111112f;  // 1/9
  p = p * z + 0.142857149f;  // 1/7
  p = p * z + 0.2f;
  p = p * z + 0.333333343f;
  p = p * z + 1.0f;
  const float lm = (2.0f * s) * p;
  const float fe = (float)e;
  return fe * 0.693145751953125f + (lm + fe * 1.42860677e-06f);
}

// d_log for a positive, normal, finite x: the same operations without the
// special-case tests (identical results wherever both apply).
MP_DM float d_log_pos(float x) {
#pragma clang fp contract(off)
  const uint32_t b = dm_float_to_bits(x);
  int e = (int)((b >> 23) & 0xffu) - 127;
  float m = dm_bits_to_float((b & 0x7fffffu) | 0x3f800000u);  // [1, 2)
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  const float s = div_core(f, 2.0f + f);
  const float z = s * s;
  float p = 0.111111112f;  // 1/9
  p = p * z + 0.142857149f;  // 1/7
  p = p * z + 0.2f;
  p = p * z + 0.333333343f;
  p = p * z + 1.0f;
  const float lm = (2.0f * s) * p;
  const float fe = (float)e;
  return fe * 0.693145751953125f + (lm + fe * 1.42860677e-06f);
}

MP_DM float d_log1p(float x) {
#pragma clang fp contract(off)
  if (x == 0.0f) return x;
  const float u = 1.0f + x;
  if (u == 1.0f) return x;
  const float d = u - 1.0f;
  return d_log(u) * (x / d);
}

MP_DM float d_exp(float x) {
#pragma clang fp contract(off)
  if (x != x) return x;
  if (x > 88.7228394f) return __builtin_inff();
  if (x < -87.3365479f) return 0.0f;  // results below FLT_MIN are flushed to 0
  const float k = __builtin_rintf(x * 1.44269502f);
  const float r = (x - k * 0.693145751953125f) - k * 1.42860677e-06f;
  float p = 1.98412701e-04f;  // 1/5040
  p = p * r + 1.38888892e-03f;  // 1/720
  p = p * r + 8.33333377e-03f;  // 1/120
  p = p * r + 4.16666679e-02f;  // 1/24
  p = p * r + 0.166666672f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  const int ki = (int)k;
  if (ki < -126) return (p * dm_pow2(ki + 64)) * 5.42101086e-20f;  // 2^-64
  if (ki > 127) return (p * dm_pow2(ki - 64)) * 1.84467441e+19f;   // 2^64
  return p * dm_pow2(ki);
}

MP_DM float d_expm1(float x) {
#pragma clang fp contract(off)
  if (x == 0.0f) return x;
  const float u = d_exp(x);
  if (u == 1.0f) return x;
  const float um1 = u - 1.0f;
  if (um1 == -1.0f) return -1.0f;
  if (u == __builtin_inff()) return u;
  return um1 * (x / d_log(u));
}

MP_DM float d_tanh(float x) {
#pragma clang fp contract(off)
  const float ax = x < 0.0f ? -x : x;
  const float t = d_expm1(-2.0f * ax);
  const float r = -t / (t + 2.0f);
  return x < 0.0f ? -r : r;
}

// cos(x) for |x| <= 2^16 (Cody-Waite with pi/2 = P1 + P2 + P3).
MP_DM float d_cos(float x) {
#pragma clang fp contract(off)
  const float k = __builtin_rintf(x * 0.636619747f);
  float r = x - k * 1.5703125f;
  r = r - k * 4.83751297e-04f;
  r = r - k * 7.54978942e-08f;
  const float z = r * r;
  // sin(r) and cos(r) Taylor polynomials, |r| <= pi/4
  float sp = -2.50521079e-08f;  // -1/11!
  sp = sp * z + 2.75573188e-06f;
  sp = sp * z + -1.98412701e-04f;
  sp = sp * z + 8.33333377e-03f;
  sp = sp * z + -0.166666672f;
  const float sn = r + (r * z) * sp;
  float cp = 2.08767559e-09f;  // 1/12!
  cp = cp * z + -2.75573188e-07f;
  cp = cp * z + 2.48015876e-05f;
  cp = cp * z + -1.38888892e-03f;
  cp = cp * z + 4.16666679e-02f;
  cp = cp * z + -0.5f;
  const float cs = 1.0f + z * cp;
  const int q = ((int)k) & 3;
  return q == 0 ? cs : q == 1 ? -sn : q == 2 ? -cs : sn;
}

MP_DM float d_pow(float b, float e) { return d_exp(e * d_log(b)); }

// ---- float64 ----

MP_DM double dd_log(double x) {
#pragma clang fp contract(off)
  if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) return x;
  union {
    double d;
    uint64_t u;
  } c;
  int e = 0;
  if (x < 2.2250738585072014e-308) {
    x = x * 4503599627370496.0;  // 2^52
    e = -52;
  }
  c.d = x;
  e += (int)((c.u >> 52) & 0x7ffu) - 1023;
  c.u = (c.u & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m = c.d;
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 23.0;
  p = p * z + 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  p = p * z + 1.0;
  const double lm = (2.0 * s) * p;
  const double de = (double)e;
  return de * 6.93147180369123816490e-01 + (lm + de * 1.90821492927058770002e-10);
}

// exp(x) in float64: Cody-Waite reduction x = n ln2 + r (split ln2, |r| <=
// ln2/2), Taylor polynomial to r^13/13! (truncation < 2^-55 relative), exact
// scaling by 2^n in two factors so that subnormal results round once.
MP_DM double dd_exp(double x) {
#pragma clang fp contract(off)
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.1332191019412) return 0.0;
  const double n = __builtin_rint(x * 1.4426950408889634);
  const double r = (x - n * 6.93147180369123816490e-01) - n * 1.90821492927058770002e-10;
  double p = 1.0 / 6227020800.0;
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
  p = p * r + 1.0;
  p = p * r + 1.0;
  const int k = (int)n;
  if (k < -1000) return __builtin_ldexp(__builtin_ldexp(p, k + 600), -600);
  return __builtin_ldexp(p, k);
}

// log Gamma(x), x > 0.
MP_DM double dd_lgamma(double x) {
#pragma clang fp contract(off)
  if (!(x > 0.0)) return __builtin_nan("");
  double shift = 0.0;
  double prod = 1.0;
  while (x < 12.0) {
    prod = prod * x;
    x = x + 1.0;
    if (prod > 1e280) {
      shift = shift + dd_log(prod);
      prod = 1.0;
    }
  }
  shift = shift + dd_log(prod);
  const double ix = 1.0 / x, ix2 = ix * ix;
  double s = 1.0 / 1188.0;  // B10 / (10 * 9)
  s = s * ix2 - 1.0 / 1680.0;
  s = s * ix2 + 1.0 / 1260.0;
  s = s * ix2 - 1.0 / 360.0;
  s = s * ix2 + 1.0 / 12.0;
  return ((x - 0.5) * dd_log(x) - x + 0.91893853320467274178) + s * ix - shift;
}

}  // namespace mp
