// Wavefront (64-lane) and workgroup reductions / scans for gfx950.
#pragma once

#include "common.hpp"

namespace mp {

constexpr int WAVE = 64;
// Bytes of the LDS reduction scratch (`red`) the block reductions may use.
constexpr int kRedUsable = 4088;

// max that propagates NaN, as jnp.max does.
__device__ __forceinline__ float nanmax(float a, float b) { return (a > b || a != a) ? a : b; }
__device__ __forceinline__ float nanmin(float a, float b) { return (a < b || a != a) ? a : b; }

// v from the lane K places lower in the same 16-lane row; the first K lanes
// of the row keep their own v (DPP row_shr:K, no LDS traffic).
template <int K>
__device__ __forceinline__ float row_shr_keep(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x110 + K, 0xf, 0xf, false));
}

// Maximum (MAX) or minimum of v over the 64 lanes of a wave, NaN if any lane
// holds NaN (the value nanmax / nanmin give in any order; of two equal zeros
// either sign may come back).  Row extremes by DPP row shifts, then the four
// row results by readlanes.  All 64 lanes must be active.
template <bool MAX>
__device__ __forceinline__ float wave_nan_extreme(float v) {
  const bool any_nan = __builtin_amdgcn_ballot_w64(v != v) != 0;
  auto op = [](float a, float b) { return MAX ? fmaxf(a, b) : fminf(a, b); };
  v = op(v, row_shr_keep<1>(v));
  v = op(v, row_shr_keep<2>(v));
  v = op(v, row_shr_keep<4>(v));
  v = op(v, row_shr_keep<8>(v));
  auto row_end = [&](int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); };
  const float r = op(op(row_end(15), row_end(31)), op(row_end(47), row_end(63)));
  return any_nan ? __builtin_nanf("") : r;
}
__device__ __forceinline__ float wave_nanmax(float v) { return wave_nan_extreme<true>(v); }
__device__ __forceinline__ float wave_nanmin(float v) { return wave_nan_extreme<false>(v); }
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// v from the lane K places lower in the same 16-lane row, 0 for the first K
// lanes of the row (DPP row_shr:K; no LDS traffic).
template <int K>
__device__ __forceinline__ uint64_t row_shr_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_update_dpp(0u, (uint32_t)v, 0x110 + K, 0xf, 0xf, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0u, (uint32_t)(v >> 32), 0x110 + K, 0xf, 0xf, false);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive prefix sum across the 64 lanes of a wave: Hillis-Steele within
// each 16-lane row by DPP row shifts, then the row totals (read from lanes 15,
// 31, 47) added to the later rows.  Integer sums: exact in any order.  All 64
// lanes must be active.
__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
  const int lane = threadIdx.x & (WAVE - 1);
  v += row_shr_u64<1>(v);
  v += row_shr_u64<2>(v);
  v += row_shr_u64<4>(v);
  v += row_shr_u64<8>(v);
  const uint64_t r0 = lane_u64(v, 15);
  const uint64_t r01 = r0 + lane_u64(v, 31);
  const uint64_t r012 = r01 + lane_u64(v, 47);
  const int row = lane >> 4;
  return v + (row == 0 ? 0ull : row == 1 ? r0 : row == 2 ? r01 : r012);
}

// #{i < n : C[i] <= t} for non-decreasing C and n >= 1: a branch-free
// search whose trip count depends on n only, so every lane of a wave runs the
// same ceil(log2 n) probes and the loop counter stays scalar.
__device__ __forceinline__ int upper_bound_fixed(const uint64_t* C, int n, uint64_t t) {
  int base = 0;
  while (n > 1) {
    const int half = n >> 1;
    base = C[base + half] <= t ? base + half : base;
    n -= half;
  }
  return base + (C[base] <= t ? 1 : 0);
}

// First index in [lo, hi) whose entry exceeds t (C non-decreasing), found by one
// whole wave: each round probes 64 evenly spaced entries and keeps the gap in
// which C crosses t, so J = 10^6 needs 4 dependent loads instead of 20.  Every
// lane must call with the same arguments; all get the result.
__device__ inline int64_t wave_upper_bound_u64(const uint64_t* C, int64_t lo, int64_t hi, uint64_t t) {
  const int lane = threadIdx.x & (WAVE - 1);
  while (hi - lo > WAVE) {
    const int64_t n = hi - lo;
    const int64_t step = (n + WAVE - 1) / WAVE;
    const int64_t off = (int64_t)(lane + 1) * step;
    const int64_t q = lo + (off < n ? off : n) - 1;
    const int c = __popcll(__ballot(C[q] <= t));  // probes are monotone: a prefix of lanes
    const int64_t cs = (int64_t)c * step;
    if (cs >= n) return hi;
    const int64_t cs1 = cs + step;
    hi = lo + (cs1 < n ? cs1 : n);
    lo = lo + cs;
  }
  const int64_t p = lo + lane;
  const bool le = p < hi && C[p] <= t;
  return lo + __popcll(__ballot(le));
}

// A key every lane holds the same value of, moved to scalar registers, so the
// arithmetic on it (a uniform threefry) runs on the scalar unit.
__device__ __forceinline__ Key uniform_key(uint32_t k0, uint32_t k1) {
  return Key{(uint32_t)__builtin_amdgcn_readfirstlane(k0), (uint32_t)__builtin_amdgcn_readfirstlane(k1)};
}

// Workgroup max (NaN-propagating) to every thread.  `red` as below.
__device__ __forceinline__ float block_nanmax(float mx, double* red) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  mx = wave_nanmax(mx);
  __syncthreads();
  float* r = reinterpret_cast<float*>(red);
  if (lane == 0) r[w] = mx;
  __syncthreads();
  mx = r[0];
  for (int i = 1; i < nw; ++i) mx = nanmax(mx, r[i]);
  return mx;
}

// Workgroup reductions.  `red` is LDS scratch of at least 2*nwaves doubles.
// All threads of the block must call; result is returned to every thread.
__device__ __forceinline__ void block_maxmin(float& mx, float& mn, double* red) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  mx = wave_nanmax(mx);
  mn = wave_nanmin(mn);
  __syncthreads();
  float* r = reinterpret_cast<float*>(red);
  if (lane == 0) { r[w] = mx; r[nw + w] = mn; }
  __syncthreads();
  mx = r[0];
  mn = r[nw];
  for (int i = 1; i < nw; ++i) { mx = nanmax(mx, r[i]); mn = nanmin(mn, r[nw + i]); }
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];  // fixed order: deterministic
  return s;
}

// Workgroup sums of N doubles per thread, in passes of as many as fit in the
// reduction scratch; each sum is taken over the waves in a fixed order.
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double* red) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  const int cap = (kRedUsable / 8) / nw;
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
  for (int k0 = 0; k0 < N; k0 += cap) {
    __syncthreads();
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < N; ++k)
        if (k >= k0 && k < k0 + cap) red[(k - k0) * nw + w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (k < k0 || k >= k0 + cap) continue;
      double s = 0.0;
      for (int i = 0; i < nw; ++i) s += red[(k - k0) * nw + i];
      v[k] = s;
    }
  }
}

}  // namespace mp
