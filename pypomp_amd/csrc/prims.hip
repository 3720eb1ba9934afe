// Device primitives behind the C ABI: the JAX key tree draws and the
// standalone systematic resampler (one workgroup per row of weights).
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

}  // namespace mp
