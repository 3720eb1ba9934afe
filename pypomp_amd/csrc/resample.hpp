// Workgroup-level systematic resampling: fixed-point prefix sums of the
// weights and the comb search (order-independent definition in common.hpp).
#pragma once

#include "blockops.hpp"
#include "engine.hpp"

namespace mp {

constexpr int kRedBytes = 4096;  // reductions (kRedUsable bytes) and two uniforms
constexpr int kUresOffset = kRedUsable;

// LDS bytes of the scan table of block_fixed_scan: rows * waves + 1 entries.
__host__ __device__ inline int64_t batched_tab_bytes(int J, int block) {
  int64_t rows = (J + block - 1) / block, nw = (block + 63) / 64;
  return align256((rows * nw + 1) * 8);
}

// Prefix sums of fixed-point weights over a workgroup in segmented form.
// Particle j = r * blockDim + tid belongs to segment k = j / 64 (the wave's 64
// lanes of row r; blockDim is a multiple of 64).  On return C[j] holds the
// inclusive sum within its segment, tab[k] the segment's exclusive offset and
// tab[n] the total, n = rows * waves: the global inclusive prefix sum of j is
// tab[j / 64] + C[j].  Wave scans in registers, one wave scans the table of
// segment totals; no pass over the particles after the table scan.
__device__ inline uint64_t block_fixed_scan_seg(const float* lw, uint64_t* C, int J, float m, uint64_t* tab) {
  const int B = blockDim.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nw = (B + 63) >> 6;
  const int rows = (J + B - 1) / B;
  const int sh = shift_for(J);
  for (int r = 0; r < rows; ++r) {
    const int j = r * B + tid;
    uint64_t e = j < J ? particle_fixed_weight(lw[j], m, sh) : 0;
    uint64_t incl = wave_incl_scan_u64(e);
    if (j < J) C[j] = incl;
    if (lane == 63) tab[r * nw + w] = incl;
  }
  __syncthreads();
  const int n = rows * nw;
  if (w == 0) {
    const int per = (n + 63) / 64;
    const int b0 = lane * per, b1 = min(n, b0 + per);
    uint64_t s = 0;
    for (int i = b0; i < b1; ++i) s += tab[i];
    uint64_t incl = wave_incl_scan_u64(s);
    uint64_t run = incl - s;
    for (int i = b0; i < b1; ++i) {
      uint64_t v = tab[i];
      tab[i] = run;
      run += v;
    }
    if (lane == 63) tab[n] = incl;
  }
  __syncthreads();
  return tab[n];
}

// Ancestor of particle j from the segmented sums of block_fixed_scan_seg:
// the last segment whose offset is <= T_j (offsets are non-decreasing and
// tab[0] = 0), then the count within it.  Equals upper_bound over the global
// prefix sums, clamped to J-1 like the reference's gather.
__device__ __forceinline__ int ancestor_seg(const uint64_t* C, const uint64_t* tab, int J, float u, uint64_t total,
                                            int j) {
  const uint64_t tj = comb_threshold(u, j, J, total);
  const int B = blockDim.x;
  const int n = ((J + B - 1) / B) * (B >> 6);
  const int k = upper_bound_fixed(tab, n, tj) - 1;
  const int start = k * 64;
  int a = start;
  if (start < J) {
    const int len = J - start < 64 ? J - start : 64;
    a += upper_bound_fixed(C + start, len, tj - tab[k]);
  }
  return a < J ? a : J - 1;
}

}  // namespace mp
