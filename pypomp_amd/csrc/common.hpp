// Shared device/host building blocks of the particle-filter engine (gfx950).
//
//  * Threefry-2x32-20 and the JAX partitionable key tree used by the reference
//    for every draw (pypomp/core/algorithms/pfilter.py:23,105,160;
//    core/model_mechanics.py:586,904; core/algorithms/helpers.py:28):
//      split(k, n)[i]       = threefry(k, (i >> 32, i & 0xffffffff))
//      random_bits(k)[i]    = w0 ^ w1 of threefry(k, (hi_i, lo_i))
//      uniform              = bitcast((bits >> 9) | 0x3f800000) - 1
//      normal               = sqrt(2) * erfinv(uniform on [nextafter(-1,0), 1))
//    erfinv is M. Giles' single-precision approximation ("Approximating the
//    erfinv function", GPU Computing Gems Jade, 2011), w = -log((1-x)(1+x)).
//  * The order-independent systematic resampler: deterministic exp, 64-bit
//    fixed-point weights and the exact comb threshold (see exp_det / fixed_weight
//    / comb_threshold below).  Every operation there is a correctly rounded IEEE
//    operation or integer arithmetic, so host and device produce the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "detmath.hpp"

#define MP_HD __host__ __device__ __forceinline__

namespace mp {

struct Key {
  uint32_t k0, k1;
};

MP_HD uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// Threefry-2x32, 20 rounds (Salmon et al., SC'11).
MP_HD void threefry2x32(uint32_t k0, uint32_t k1, uint32_t& x0, uint32_t& x1) {
  const uint32_t k2 = k0 ^ k1 ^ 0x1BD11BDAu;
  x0 += k0;
  x1 += k1;
#define MP_TF_ROUND(r) \
  x0 += x1;            \
  x1 = rotl32(x1, r);  \
  x1 ^= x0;
  MP_TF_ROUND(13) MP_TF_ROUND(15) MP_TF_ROUND(26) MP_TF_ROUND(6)
  x0 += k1; x1 += k2 + 1u;
  MP_TF_ROUND(17) MP_TF_ROUND(29) MP_TF_ROUND(16) MP_TF_ROUND(24)
  x0 += k2; x1 += k0 + 2u;
  MP_TF_ROUND(13) MP_TF_ROUND(15) MP_TF_ROUND(26) MP_TF_ROUND(6)
  x0 += k0; x1 += k1 + 3u;
  MP_TF_ROUND(17) MP_TF_ROUND(29) MP_TF_ROUND(16) MP_TF_ROUND(24)
  x0 += k1; x1 += k2 + 4u;
  MP_TF_ROUND(13) MP_TF_ROUND(15) MP_TF_ROUND(26) MP_TF_ROUND(6)
  x0 += k2; x1 += k0 + 5u;
#undef MP_TF_ROUND
}

// split(k, n)[i] -- independent of n (partitionable mode).
MP_HD Key split_at(Key k, uint64_t i) {
  uint32_t x0 = (uint32_t)(i >> 32), x1 = (uint32_t)i;
  threefry2x32(k.k0, k.k1, x0, x1);
  return Key{x0, x1};
}

// 32 random bits at flat index i of a draw of any shape.
MP_HD uint32_t bits_at(Key k, uint64_t i) {
  uint32_t x0 = (uint32_t)(i >> 32), x1 = (uint32_t)i;
  threefry2x32(k.k0, k.k1, x0, x1);
  return x0 ^ x1;
}

MP_HD float unit_float(uint32_t bits) {
  return __uint_as_float((bits >> 9) | 0x3F800000u) - 1.0f;
}

// jax.random.uniform(k, ..., minval=0, maxval=1) from its bits.
MP_HD float uniform01(uint32_t bits) {
  float f = unit_float(bits);
  return f > 0.0f ? f : 0.0f;
}

// Giles' single-precision erfinv with his w = -log((1-x)(1+x)) (deterministic
// log, see detmath.hpp).  The tail branch (|x| > 0.9966, w >= 5) is entered only
// by waves that hold such a lane.
// |x| < 1 gives (1-x)(1+x) >= 2^-24 (x = nextafter(-1, 0) at the extreme),
// a positive normal number; |x| == 1 is overridden below.
__device__ __forceinline__ float erfinv_f32(float x) {
  const float w = -d_log_pos((1.0f - x) * (1.0f + x));
  const float wc = w - 2.5f;
  float p = 2.81022636e-08f;
  p = 3.43273939e-07f + p * wc;
  p = -3.5233877e-06f + p * wc;
  p = -4.39150654e-06f + p * wc;
  p = 0.00021858087f + p * wc;
  p = -0.00125372503f + p * wc;
  p = -0.00417768164f + p * wc;
  p = 0.246640727f + p * wc;
  p = 1.50140941f + p * wc;
  const bool tail = !(w < 5.0f);
  if (__builtin_amdgcn_ballot_w64(tail)) {
    const float wt = sqrtf(w) - 3.0f;
    float q = -0.000200214257f;
    q = 0.000100950558f + q * wt;
    q = 0.00134934322f + q * wt;
    q = -0.00367342844f + q * wt;
    q = 0.00573950773f + q * wt;
    q = -0.0076224613f + q * wt;
    q = 0.00943887047f + q * wt;
    q = 1.00167406f + q * wt;
    q = 2.83297682f + q * wt;
    if (tail) p = q;
  }
  float r = p * x;
  if (fabsf(x) == 1.0f) r = x * 3.40282347e+38f;
  return r;
}

// jax.random.normal from 32 random bits.
__device__ __forceinline__ float normal_from_bits(uint32_t bits) {
  const float lo = -0.99999994f;  // nextafter(-1, 0)
  // (hi - lo) rounds to 2.0f in float32; f * 2 is exact.
  float u = unit_float(bits) * 2.0f + lo;
  u = u > lo ? u : lo;
  return 1.41421354f * erfinv_f32(u);
}

__device__ __forceinline__ float normal_at(Key k, uint64_t i) { return normal_from_bits(bits_at(k, i)); }

// ---------------------------------------------------------------------------
// Deterministic resampling arithmetic.
//
// exp_det(x), x <= 0: k = rint(x*log2e); r = (x - k*ln2_hi) - k*ln2_lo;
// p = Horner of sum_{n<=13} r^n/n! with separate (uncontracted) multiplies and
// adds; value = ldexp(p, k).  Every step is correctly rounded IEEE double
// arithmetic, so any conforming host or device computes identical bits.
// ---------------------------------------------------------------------------
#define MP_EXPC13 0x1.6124613a86d09p-33
#define MP_EXPC12 0x1.1eed8eff8d898p-29
#define MP_EXPC11 0x1.ae64567f544e4p-26
#define MP_EXPC10 0x1.27e4fb7789f5cp-22
#define MP_EXPC9 0x1.71de3a556c734p-19
#define MP_EXPC8 0x1.a01a01a01a01ap-16
#define MP_EXPC7 0x1.a01a01a01a01ap-13
#define MP_EXPC6 0x1.6c16c16c16c17p-10
#define MP_EXPC5 0x1.1111111111111p-7
#define MP_EXPC4 0x1.5555555555555p-5
#define MP_EXPC3 0x1.5555555555555p-3
#define MP_EXPC2 0x1p-1
#define MP_EXPC1 0x1p+0
#define MP_EXPC0 0x1p+0

MP_HD int shift_for(int64_t J) {
  int b = 0;
  uint64_t v = (uint64_t)J;
  while (v) { ++b; v >>= 1; }
  return 62 - b;
}

// Fixed-point weight floor(exp(d) * 2^sh) for d = lw - max(lw) <= 0 (or NaN);
// sh = 62 - bitlen(J) keeps every prefix sum below 2^62.
MP_HD uint64_t fixed_weight(double d, int sh) {
#pragma clang fp contract(off)
  const double cutoff = -(double)(sh + 2) * 0.6931471805599453;
  if (!(d >= cutoff)) return 0;  // also NaN
  const double LOG2E = 1.4426950408889634;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  double k = rint(d * LOG2E);
  double r = d - k * LN2_HI;
  r = r - k * LN2_LO;
  double p = MP_EXPC13;
  p = p * r + MP_EXPC12;
  p = p * r + MP_EXPC11;
  p = p * r + MP_EXPC10;
  p = p * r + MP_EXPC9;
  p = p * r + MP_EXPC8;
  p = p * r + MP_EXPC7;
  p = p * r + MP_EXPC6;
  p = p * r + MP_EXPC5;
  p = p * r + MP_EXPC4;
  p = p * r + MP_EXPC3;
  p = p * r + MP_EXPC2;
  p = p * r + MP_EXPC1;
  p = p * r + MP_EXPC0;
  double v = ldexp(p, (int)k + sh);
  return (uint64_t)v;  // v >= 0 and < 2^62: truncation is the floor
}

// Weight of particle with log-weight lw given the filter's max log-weight m.
// m == -inf (all -inf/NaN): uniform weight 1.  m == +inf: 1 for +inf particles.
MP_HD uint64_t particle_fixed_weight(float lw, float m, int sh) {
  if (m == -__builtin_inff()) return 1;
  if (m == __builtin_inff()) return lw == __builtin_inff() ? 1 : 0;
  return fixed_weight((double)lw - (double)m, sh);
}

// Log-likelihood term of one observation, m + log(sum_j exp(w_j - m)), with the
// sum taken as the exact integer total of the fixed-point weights times 2^-sh
// (the weights the resampler uses).  Integer totals do not depend on the
// summation order, and dd_log is a fixed IEEE sequence, so every launch shape,
// both execution forms and the oracle get the same bits.  The truncation of
// each weight to 2^-sh of the largest costs < J 2^-sh relative (J = 10^6:
// 5e-7), the size of the reference's float32 sum error.  m == -inf: -inf
// (every weight zero); m == +inf or NaN: NaN, as exp(inf - inf) makes it in the
// reference (_normalize_weights, core/algorithms/helpers.py:34-57).
MP_HD double step_loglik(float m, uint64_t total, int sh) {
#pragma clang fp contract(off)
  if (m == -__builtin_inff()) return -__builtin_inf();
  if (!(m < __builtin_inff())) return __builtin_nan("");
  return (double)m + dd_log(ldexp((double)total, -sh));
}

// T_j = floor(unif_j * total) with unif_j = fl32(fl32(u + j) / J), exactly.
// The division: u + j is 0 or at least 2^-23 and J < 2^31, so the quotient
// stays far above underflow and div_by is exact (1 / J is the same for every
// particle and leaves the loops).  0 <= j < J < 2^31: j converts to float as
// a 32-bit integer (the same correctly rounded value, one instruction).
MP_HD uint64_t comb_threshold(float u, int64_t j, int64_t J, uint64_t total) {
  const float fJ = (float)J;
  float f = div_by(u + (float)(uint32_t)j, fJ, 1.0f / fJ);
  uint32_t b = __float_as_uint(f);
  uint32_t e = (b >> 23) & 0xffu;
  if (e == 0) return 0;  // f == 0 (subnormal f cannot occur: f >= 2^-23 / J)
  uint64_t mant = (uint64_t)((b & 0x7fffffu) | 0x800000u);
  int sh = 150 - (int)e;  // f = mant * 2^-sh, sh >= 23 since f <= 1
  uint64_t p_lo = mant * (total & 0xffffffffull);
  uint64_t p_hi = mant * (total >> 32);
  uint64_t X = p_hi + (p_lo >> 32);
  uint64_t lo32 = p_lo & 0xffffffffull;
  if (sh >= 32) {
    int s = sh - 32;
    return s >= 64 ? 0 : (X >> s);
  }
  return (X << (32 - sh)) + (lo32 >> sh);
}

// #{i < n : C[i] <= t} for a non-decreasing array C (upper bound).
template <class Ptr>
MP_HD int64_t upper_bound_u64(Ptr C, int64_t lo, int64_t hi, uint64_t t) {
  while (lo < hi) {
    int64_t mid = lo + ((hi - lo) >> 1);
    if (C[mid] <= t) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

}  // namespace mp
