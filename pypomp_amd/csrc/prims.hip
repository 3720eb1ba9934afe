// Device primitives behind the C ABI: the JAX key tree draws and the
// standalone systematic resampler (one workgroup per row of weights).
#include "engine.hpp"
#include "resample.hpp"

namespace mp {

__global__ void k_bits(const uint32_t* __restrict__ keys, int64_t nkeys, int64_t n, uint32_t* __restrict__ out,
                       int mode) {
  const int64_t total = nkeys * n;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = g / n, i = g - k * n;
    const Key key{keys[2 * k], keys[2 * k + 1]};
    if (mode == 3) {
      Key s = split_at(key, (uint64_t)i);
      out[2 * g] = s.k0;
      out[2 * g + 1] = s.k1;
      continue;
    }
    const uint32_t b = bits_at(key, (uint64_t)i);
    if (mode == 0) out[g] = b;
    else if (mode == 1) reinterpret_cast<float*>(out)[g] = uniform01(b);
    else reinterpret_cast<float*>(out)[g] = normal_from_bits(b);
  }
}

hipError_t launch_bits(const uint32_t* keys, int64_t nkeys, int64_t n, void* out, int mode, hipStream_t s) {
  const int64_t total = nkeys * n;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_bits, dim3((unsigned)blocks), dim3(256), 0, s, keys, nkeys, n, (uint32_t*)out, mode);
  return hipGetLastError();
}

__global__ __launch_bounds__(1024) void k_resample_rows(const float* __restrict__ logw, const uint32_t* __restrict__ keys,
                                                        int64_t R, int J, int32_t* __restrict__ anc,
                                                        uint64_t* __restrict__ Cws) {
  extern __shared__ __align__(256) char smem[];
  double* red = reinterpret_cast<double*>(smem);
  uint64_t* tab = reinterpret_cast<uint64_t*>(smem + kRedBytes);
  for (int64_t r = blockIdx.x; r < R; r += gridDim.x) {
    const float* lw = logw + r * J;
    uint64_t* C = Cws + r * J;
    float mx = -__builtin_inff(), mn = __builtin_inff();
    for (int j = threadIdx.x; j < J; j += blockDim.x) {
      float v = lw[j];
      if (v == v) mx = fmaxf(mx, v);  // NaN weights carry no mass
    }
    block_maxmin(mx, mn, red);
    const uint64_t total = block_fixed_scan_seg(lw, C, J, mx, tab);
    const Key k{keys[2 * r], keys[2 * r + 1]};
    const float u = uniform01(bits_at(k, 0));
    for (int j = threadIdx.x; j < J; j += blockDim.x) anc[r * J + j] = ancestor_seg(C, tab, J, u, total, j);
    __syncthreads();
  }
}

hipError_t launch_resample_rows(const float* logw, const uint32_t* keys, int64_t R, int J, int32_t* anc,
                                uint64_t* C, hipStream_t s) {
  const int block = 256;
  const size_t lds = kRedBytes + batched_tab_bytes(J, block);
  int64_t grid = R < 4096 ? R : 4096;
  hipLaunchKernelGGL(k_resample_rows, dim3((unsigned)grid), dim3(block), lds, s, logw, keys, R, J, anc, C);
  return hipGetLastError();
}

// VALU issue calibration (no counterpart in the reference): every lane runs
// `iters` rounds of eight independent dependency chains, so the issue rate and
// not latency bounds the kernel.  kind 0: float FMAs (v_fma_f32, no packed
// forms); 1: the integer add / rotate / xor rounds of threefry; 2: FMAs
// interleaved with base-2 exponentials (the transcendental unit); 3: FMAs
// interleaved with integer add / xor; 4: integer add / xor alone.  The result
// is written so that nothing is optimised away.
template <int KIND>
__global__ __launch_bounds__(256) void k_valu_probe(int64_t iters, float* out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (KIND == 3 || KIND == 4) {
    // 3: float FMA chains interleaved with integer add / xor chains;
    // 4: integer add / xor chains alone
    uint32_t u[8];
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      u[k] = (uint32_t)g * 2654435761u + (uint32_t)k;
      x[k] = 1.0f + 1e-7f * (float)(g + k);
    }
    const float b = 0.999999f, c = 1e-7f;
    for (int64_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if constexpr (KIND == 3) x[k] = __builtin_fmaf(x[k], b, c);
        u[k] = (u[k] + 0x9e3779b9u) ^ (uint32_t)k;
      }
    }
    uint32_t a = 0;
    float f = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a ^= u[k];
      f += x[k];
    }
    out[g] = f + __uint_as_float(a & 0x3fffffffu);
  } else if constexpr (KIND == 1) {
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (uint32_t)g * 2654435761u + (uint32_t)k;
    for (int64_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        uint32_t y = x[k] + 0x9e3779b9u;
        y = rotl32(y, 13);
        x[k] = y ^ x[(k + 1) & 7];
      }
    }
    uint32_t a = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) a ^= x[k];
    out[g] = __uint_as_float(a & 0x3fffffffu);
  } else {
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = 1.0f + 1e-7f * (float)(g + k);
    const float b = 0.999999f, c = 1e-7f;
    for (int64_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        x[k] = __builtin_fmaf(x[k], b, c);
        if constexpr (KIND == 2) x[k] = __builtin_amdgcn_exp2f(x[k]);
      }
    }
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) a += x[k];
    out[g] = a;
  }
}

hipError_t launch_valu_probe(int kind, int64_t iters, int blocks, float* out, hipStream_t s) {
  if (kind == 0) mp_launch(k_valu_probe<0>, dim3((unsigned)blocks), dim3(256), 0, s, iters, out);
  else if (kind == 1) mp_launch(k_valu_probe<1>, dim3((unsigned)blocks), dim3(256), 0, s, iters, out);
  else if (kind == 2) mp_launch(k_valu_probe<2>, dim3((unsigned)blocks), dim3(256), 0, s, iters, out);
  else if (kind == 3) mp_launch(k_valu_probe<3>, dim3((unsigned)blocks), dim3(256), 0, s, iters, out);
  else mp_launch(k_valu_probe<4>, dim3((unsigned)blocks), dim3(256), 0, s, iters, out);
  return hipGetLastError();
}

}  // namespace mp
