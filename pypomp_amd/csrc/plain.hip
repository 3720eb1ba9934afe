// Plain form: one filter at a time, its particles spread over the whole GPU,
// two launches per observation (the shape of the reference's per-step XLA
// loop body, pfilter.py:97-182):
//
//  k_plain_prop  -- ancestor search (systematic comb against the previous
//                   step's prefix sums) fused with the gather, the Euler
//                   sub-steps, the measurement density and the log-weight;
//                   per-tile max/min, the last tile combines them.
//  k_plain_scan  -- fixed-point weights, per-tile exp-sums and one-pass
//                   inclusive prefix sum across tiles with decoupled look-back;
//                   the last tile finishes the log-likelihood term (FP64 sum
//                   over t), the resampling decision and the comb's uniform.
#include "blockops.hpp"
#include "engine.hpp"
#include "seir.hpp"

namespace mp {

constexpr unsigned kAgg = 1, kIncl = 2;

__device__ __forceinline__ unsigned atomic_ticket(unsigned int* p) {
  return __hip_atomic_fetch_add(p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}

template <class M>
__global__ __launch_bounds__(256) void k_plain_prop(const PlainArgs a) {
  __shared__ double red[32];
  __shared__ int is_last;
  const ModelDesc md = a.md;
  const int J = a.J, S = M::FIXED_S ? M::MAXS : md.S, t = a.t, tid = threadIdx.x, B = blockDim.x;
  const int64_t o0 = a.ser.obs_off[a.unit];
  const int ti0 = a.ser.tidx0[o0 + t];
  const int nst = a.ser.nstep[o0 + t];
  const float* dts = a.ser.dt_ext + a.ser.dt_off[a.unit];
  const float ty = a.ser.times[o0 + t];
  const float tstart = t == 0 ? a.ser.t0 : a.ser.times[o0 + t - 1];
  const float* y = a.ser.ys + (o0 + t) * md.dy;
  const int64_t base = ti0 + (int64_t)t;
  const float fJ = (float)J;
  const float logJ = d_log(fJ);

  float th[M::MAXP];
  MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < md.d) th[p] = a.theta[p];
  if (a.flags & F_SHOULD_TRANS) M::from_est(md, th);
  typename M::P P;
  M::prepare(md, th, P);

  const PlainState st = *a.st;
  const float* Xc = (t & 1) ? a.X1 : a.X0;
  float* Xn = (t & 1) ? a.X0 : a.X1;
  const float wconst = t == 0 ? d_log(1.0f / fJ) : st.wconst;
  const bool uniform_w = t == 0 ? true : st.uniform_w;
  const bool have_anc = t == 0 ? false : st.have_anc;

  float lmax = -__builtin_inff(), lmin = __builtin_inff();
  const int64_t tile0 = (int64_t)blockIdx.x * a.rows * B;
  for (int r = 0; r < a.rows; ++r) {
    const int64_t j = tile0 + (int64_t)r * B + tid;
    if (j >= J) break;
    float x[M::MAXS];
    int64_t src = j;
    if (t == 0) {
      M::rinit(md, P, split_at(Key{a.fkey[0], a.fkey[1]}, 1 + (uint64_t)j), a.ser.t0, x);
    } else {
      if (have_anc) {
        uint64_t tj = comb_threshold(st.ures, j, J, st.total);
        int64_t an = upper_bound_u64(a.C, 0, J, tj);
        src = an < J ? an : J - 1;
      }
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = Xc[(int64_t)s * J + src];
    }
    float wprev = uniform_w ? wconst : (st.collapsed ? -logJ : a.lw[j] - st.llf);
    MP_UNROLL for (int s = 0; s < M::MAXS; ++s)
      if ((md.accum_mask >> s) & 1u) x[s] = 0.f;
    float tt = tstart;
    for (int k = 0; k < nst; ++k) {
      const Key sk{a.sched[2 * (base + k)], a.sched[2 * (base + k) + 1]};
      const float dt = dts[ti0 + k];
      M::step(md, P, split_at(sk, (uint64_t)j), tt, dt, x);
      tt = tt + dt;
    }
    const float w = wprev + M::dmeas(md, P, y, x, ty);
    MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) Xn[(int64_t)s * J + j] = x[s];
    a.lw[j] = w;
    lmax = nanmax(lmax, w);
    lmin = nanmin(lmin, w);
  }
  block_maxmin(lmax, lmin, red);
  if (tid == 0) {
    a.pmax[blockIdx.x] = lmax;
    a.pmin[blockIdx.x] = lmin;
    unsigned k = atomic_ticket(&a.st->ticket_prop);
    is_last = (k == gridDim.x - 1);
  }
  __syncthreads();
  if (is_last) {
    float mx = -__builtin_inff(), mn = __builtin_inff();
    for (int i = tid; i < (int)gridDim.x; i += B) {
      mx = nanmax(mx, __hip_atomic_load(&a.pmax[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      mn = nanmin(mn, __hip_atomic_load(&a.pmin[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    block_maxmin(mx, mn, red);
    if (tid == 0) {
      a.st->M = mx;
      a.st->mn = mn;
      a.st->ticket_prop = 0;
    }
  }
}

template <class M>
__global__ __launch_bounds__(256) void k_plain_scan(const PlainArgs a) {
  __shared__ double red[32];
  __shared__ uint64_t tab[8 * 4 + 1];
  __shared__ uint64_t s_prefix;
  __shared__ int s_tile, is_last;
  const int J = a.J, tid = threadIdx.x, B = blockDim.x, lane = tid & 63, w = tid >> 6;
  const int nw = (B + 63) >> 6;
  const float m = a.st->M;
  const int sh = shift_for(J);
  const unsigned epoch = (unsigned)a.t + 1u;

  if (tid == 0) s_tile = (int)atomic_ticket(&a.st->tile_ctr);
  __syncthreads();
  const int tile = s_tile;
  const int64_t tile0 = (int64_t)tile * a.rows * B;

  // per-thread fixed-point weights and exp sums; wave scans per row
  double se = 0.0;
  const bool live = m == m && m != -__builtin_inff();
  uint64_t rowincl[8];
  for (int r = 0; r < a.rows; ++r) {
    const int64_t j = tile0 + (int64_t)r * B + tid;
    uint64_t e = 0;
    if (j < J) {
      const float v = a.lw[j];
      e = particle_fixed_weight(v, m, sh);
      if (live && v == v) se += (double)expf(v - m);
    }
    uint64_t inc = wave_incl_scan_u64(e);
    rowincl[r] = inc;
    if (lane == 63) tab[r * nw + w] = inc;
  }
  __syncthreads();
  const int n = a.rows * nw;
  if (tid == 0) {
    uint64_t run = 0;
    for (int i = 0; i < n; ++i) {
      uint64_t v = tab[i];
      tab[i] = run;
      run += v;
    }
    tab[n] = run;  // tile aggregate
    // decoupled look-back
    const uint64_t aggv = run;
    uint64_t prefix = 0;
    if (tile == 0) {
      a.incl[0] = aggv;
      __hip_atomic_store(&a.flag[0], (epoch << 2) | kIncl, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      a.agg[tile] = aggv;
      __hip_atomic_store(&a.flag[tile], (epoch << 2) | kAgg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      int k = tile - 1;
      while (true) {
        unsigned fl = __hip_atomic_load(&a.flag[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if ((fl >> 2) != epoch) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        if ((fl & 3u) == kIncl) {
          prefix += __hip_atomic_load(&a.incl[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        prefix += __hip_atomic_load(&a.agg[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        --k;
      }
      a.incl[tile] = prefix + aggv;
      __hip_atomic_store(&a.flag[tile], (epoch << 2) | kIncl, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_prefix = prefix;
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  for (int r = 0; r < a.rows; ++r) {
    const int64_t j = tile0 + (int64_t)r * B + tid;
    if (j < J) a.C[j] = prefix + tab[r * nw + w] + rowincl[r];
  }
  se = block_sum(se, red);
  if (tid == 0) {
    a.psum[tile] = se;
    unsigned k = atomic_ticket(&a.st->ticket_scan);
    is_last = (k == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  // last tile: finish the step (deterministic fixed-order sums)
  double s = 0.0;
  for (int i = tid; i < (int)gridDim.x; i += B) s += __hip_atomic_load(&a.psum[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s = block_sum(s, red);
  if (tid == 0) {
    PlainState* st = a.st;
    const float fJ = (float)J, logJ = d_log(fJ);
    const bool collapsed = (m == -__builtin_inff());
    const double llt = collapsed ? -__builtin_inf() : (double)m + log(s);
    const float llf = (float)llt;
    st->loglik = (a.t == 0 ? 0.0 : st->loglik) + llt;
    if (a.cll) a.cll[a.t] = llf;
    const float nmax = collapsed ? -logJ : m - llf;
    const float nmin = collapsed ? -logJ : st->mn - llf;
    const bool res = (nmax - nmin) > a.logthresh;
    st->llf = llf;
    st->collapsed = collapsed;
    st->have_anc = res;
    st->uniform_w = res;
    st->wconst = 0.0f - logJ;
    st->total = __hip_atomic_load(&a.C[J - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int64_t o0 = a.ser.obs_off[a.unit];
    const int64_t base = a.ser.tidx0[o0 + a.t] + (int64_t)a.t;
    const int nst = a.ser.nstep[o0 + a.t];
    const Key rk{a.sched[2 * (base + nst)], a.sched[2 * (base + nst) + 1]};
    st->ures = uniform01(bits_at(rk, 0));
    st->ticket_scan = 0;
    st->tile_ctr = 0;
    if (a.t == a.T - 1) *a.loglik_out = st->loglik;
  }
}

// Start of a filter: clear the look-back flags and the step state.
__global__ void k_plain_reset(unsigned int* flag, int ntiles, PlainState* st) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntiles) flag[i] = 0u;
  if (i == 0) {
    PlainState z{};
    *st = z;
  }
}

hipError_t plain_reset(unsigned int* flag, int ntiles, PlainState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_plain_reset, dim3((ntiles + 255) / 256), dim3(256), 0, s, flag, ntiles, st);
  return hipGetLastError();
}

// which = 0: propagate kernel, 1: scan kernel
template <class M>
static hipError_t plain_kernel_t(int which, const PlainArgs& a, int ntiles, hipStream_t s) {
  if (which == 0) hipLaunchKernelGGL(k_plain_prop<M>, dim3(ntiles), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_plain_scan<M>, dim3(ntiles), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t plain_kernel(int which, const PlainArgs& a, int ntiles, hipStream_t s) {
  switch (a.md.kind) {
    case MODEL_LG:
      if (a.md.dx == 1 && a.md.dy == 1) return plain_kernel_t<LG1>(which, a, ntiles, s);
      if (a.md.dx == 2 && a.md.dy == 2) return plain_kernel_t<LG2>(which, a, ntiles, s);
      return plain_kernel_t<LG4>(which, a, ntiles, s);
    case MODEL_SV: return plain_kernel_t<SV>(which, a, ntiles, s);
    case MODEL_SEIR: return plain_kernel_t<SEIR>(which, a, ntiles, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace mp
