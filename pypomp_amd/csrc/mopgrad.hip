// MOP(alpha) value and hand-written adjoint gradient, one workgroup per filter.
//
// The reference differentiates _mop_internal (pypomp/core/algorithms/mop.py:19-135)
// with jax.grad and one jax.checkpoint per step (mop.py:54; ad_helpers.py:6-8).
// Here the gradient is the exact reverse-mode derivative of that estimator
// with the resampling ancestry and the random numbers held fixed:
//
//   loglik = sum_t [LSE(wP_t + m_t) - LSE(wP_t)],   wP_t = alpha wF_{t-1},
//   wF_t^i = wP_t^{a_t^i} + m_t^{a_t^i} - sg(m_t^{a_t^i}),
//   x_t^j = f(x_{t-1}^{a_{t-1}^j}, theta, z_t^j),  m_t^j = g(y_t, x_t^j, theta).
//
// Since wP_t is constant across particles in value, softmax(wP_t + m_t) = pi_t
// (normalised exp(m_t)) and softmax(wP_t) = 1/J.  With Lambda_t^j and Xc_t^j
// the sums of lambda_t^i and of the step-VJPs over the children i of
// particle j (a_t^i = j):
//   gbar_t^j   = pi_t^j + Lambda_t^j                (adjoint of m_t^j)
//   xbar_t^j   = gbar_t^j dg/dx + Xc_t^j
//   lambda_{t-1}^j = alpha (pi_t^j - 1/J + Lambda_t^j)
//   thetabar  += gbar_t^j dg/dtheta + (df/dtheta)^T xbar_t^j (+ rinit at t = 0).
// Systematic resampling makes a_t non-decreasing, so every parent's children are
// one contiguous run: the scatter is a deterministic segmented sum.
//
// Memory: the forward keeps every K-th pre-step state, the ancestry of every
// step (4 bytes per particle-step) and each step's density maximum and exp-sum;
// the reverse sweep recomputes each K-step block's states from its checkpoint
// with the logged ancestry (no reductions, scans or searches) and runs the
// block backwards, evaluating each density where its adjoint needs it.  States
// storage O((T/K + K) J S) instead of O(T J S).  The model's step VJP reads its
// noise from the stored output state.
#include "blockops.hpp"
#include "engine.hpp"
#include "resample.hpp"
#include "seir.hpp"

namespace mp {

// Per-filter global scratch: checkpoints, the ancestry log, the step-noise
// log, per-step statistics and the recomputed K-step block.
struct GradScratch {
  int64_t ckx, anc, noise, stat, bx, work, total;
};

// Per-step working arrays.  The forward set (prefix sums, log-weights, two
// state buffers) and the reverse set (children's contributions, carried
// adjoints, the step's ancestry) are never live together, so they share one region,
// kept in LDS when it fits and in the global scratch otherwise.
struct GradWork {
  int64_t c, lw, x0, x1;                 // forward
  int64_t cx, cl, xca, lama, ap;         // reverse
  int64_t total;
};

__host__ __device__ inline int64_t align16(int64_t v) { return (v + 15) & ~int64_t(15); }

__host__ __device__ inline GradWork grad_work_layout(int S, int J) {
  GradWork w;
  int64_t o = 0;
  w.c = o;    o = align16(o + (int64_t)J * 8);
  w.lw = o;   o = align16(o + (int64_t)J * 4);
  w.x0 = o;   o = align16(o + (int64_t)S * J * 4);
  w.x1 = o;   o = align16(o + (int64_t)S * J * 4);
  const int64_t fwd = o;
  o = 0;
  w.cx = o;   o = align16(o + (int64_t)S * J * 4);
  w.cl = o;   o = align16(o + (int64_t)J * 4);
  w.xca = o;  o = align16(o + (int64_t)S * J * 4);
  w.lama = o; o = align16(o + (int64_t)J * 4);
  w.ap = o;   o = align16(o + (int64_t)J * 4);
  w.total = fwd > o ? fwd : o;
  return w;
}

// nz: noise words per particle-step (the model's NOISE)
__host__ __device__ inline GradScratch grad_scratch_layout(int S, int nz, int J, int T, int K, bool lds_work) {
  GradScratch g;
  const int64_t nblk = (T + K - 1) / K;
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    int64_t r = o;
    o += align256(bytes);
    return r;
  };
  g.ckx = take(nblk * S * (int64_t)J * 4);
  g.anc = take((int64_t)T * J * 4);
  g.noise = take((int64_t)T * nz * J * 4);
  g.stat = take((int64_t)T * 16);
  g.bx = take((int64_t)K * S * J * 4);
  g.work = take(lds_work ? 0 : grad_work_layout(S, J).total);
  g.total = o;
  return g;
}

// Forward pass (value): the ancestry a_t of every step (a_0 = identity), the
// step noise of every particle-step (models whose noise depends on the
// particle's key alone: draw + step_z == step), the maximum and FP64 exp-sum
// of each step's measurement densities, and the pre-step states at every K-th
// step.  Reverse sweep, block by block: the block's states are recomputed from
// its checkpoint with the logged ancestry and noise (gather and step_z, no
// random numbers drawn again), then the adjoint runs backwards through the
// block.
template <class M>
__device__ __forceinline__ void mop_grad_filter(const GradArgs& a, const int f, char* buf, char* wbuf, double* red,
                                                uint64_t* tab) {
  const ModelDesc md = a.md;
  // the filter's panel unit: its own rows of the series, its own length; the
  // scratch is laid out for the longest unit
  const int u = a.unit ? a.unit[f] : 0;
  const int64_t o0 = a.ser.obs_off[u];
  const int T = (int)(a.ser.obs_off[u + 1] - o0);
  const int J = a.J, S = M::FIXED_S ? M::MAXS : md.S, d = md.d, K = a.K;
  const int tid = threadIdx.x, B = blockDim.x;
  constexpr int NZ = M::NOISE > 0 ? M::NOISE : 1;
  const GradScratch L = grad_scratch_layout(S, M::NOISE, J, a.ser.Tmax, K, a.lds_work);
  const GradWork W = grad_work_layout(S, J);
  float* ckx = reinterpret_cast<float*>(buf + L.ckx);
  int32_t* ancl = reinterpret_cast<int32_t*>(buf + L.anc);
  float* zlog = reinterpret_cast<float*>(buf + L.noise);  // (T, NOISE, J)
  double* stat = reinterpret_cast<double*>(buf + L.stat);
  float* bx = reinterpret_cast<float*>(buf + L.bx);
  uint64_t* Cc = reinterpret_cast<uint64_t*>(wbuf + W.c);
  float* lw = reinterpret_cast<float*>(wbuf + W.lw);
  float* cx = reinterpret_cast<float*>(wbuf + W.cx);
  float* cl = reinterpret_cast<float*>(wbuf + W.cl);
  float* xca = reinterpret_cast<float*>(wbuf + W.xca);
  float* lama = reinterpret_cast<float*>(wbuf + W.lama);
  int32_t* ap = reinterpret_cast<int32_t*>(wbuf + W.ap);  // this step's ancestry
  const int64_t SJ = (int64_t)S * J;
  const float* ys = a.ser.ys + o0 * md.dy;
  const float* times = a.ser.times + o0;
  const int32_t* tidx0 = a.ser.tidx0 + o0;
  const float* dts = a.ser.dt_ext + a.ser.dt_off[u];
  const uint32_t* sch = a.sched + 2 * (int64_t)f * a.sched_len;
  const Key fkey = uniform_key(a.keys[2 * f], a.keys[2 * f + 1]);
  const float fJ = (float)J, logJ = d_log(fJ), invJ = 1.0f / fJ;
  const int sh = shift_for(J);

  float th_est[M::MAXP], th[M::MAXP];
  MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < d) {
    th_est[p] = a.theta[f * a.theta_stride + p];
    th[p] = th_est[p];
  }
  if (!md.identity_trans) M::from_est(md, th);
  typename M::P P;
  M::prepare(md, th, P);

  auto step_key = [&](int t) -> Key {
    const int64_t base = tidx0[t] + (int64_t)t;
    return uniform_key(sch[2 * base], sch[2 * base + 1]);
  };
  auto res_key = [&](int t) -> Key {
    const int64_t base = tidx0[t] + (int64_t)t + 1;  // one sub-step per observation
    return uniform_key(sch[2 * base], sch[2 * base + 1]);
  };
  auto tstart = [&](int t) { return t == 0 ? a.ser.t0 : times[t - 1]; };

  // ---------------- forward: value, ancestry log, checkpoints ----------------
  float* XA = reinterpret_cast<float*>(wbuf + W.x0);
  float* XB = reinterpret_cast<float*>(wbuf + W.x1);
  for (int j = tid; j < J; j += B) {
    float x[M::MAXS];
    M::rinit(md, P, split_at(fkey, 1 + (uint64_t)j), a.ser.t0, x);
    MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) XA[s * (int64_t)J + j] = x[s];
  }
  __syncthreads();
  double loglik = 0.0;
  uint64_t total = 0;
  float ures = 0.f;
  // the comb's uniform of step t, drawn by the first wave while the others
  // propagate and read after the step's barriers (two slots: step t writes
  // slot t & 1 while step t-1's value is still being read from the other)
  float* ures_slot = reinterpret_cast<float*>(reinterpret_cast<char*>(red) + kUresOffset);
  for (int t = 0; t < T; ++t) {
    if (tid < WAVE && t < T - 1) {
      const float u = uniform01(bits_at(res_key(t), 0));
      if (tid == 0) ures_slot[t & 1] = u;
    }
    const bool ck = (t % K) == 0;
    float* ckt = ckx + (int64_t)(t / K) * SJ;
    int32_t* anct = ancl + (int64_t)t * J;
    const Key sk = step_key(t);
    const float dt = dts[tidx0[t]];
    const float ts = tstart(t), tt = times[t];
    const float* y = ys + (int64_t)t * md.dy;
    float lmax = -__builtin_inff();
    for (int j = tid; j < J; j += B) {
      const int src = t == 0 ? j : ancestor_seg(Cc, tab, J, ures, total, j);
      if (ck) MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) ckt[s * (int64_t)J + j] = XA[s * (int64_t)J + j];
      anct[j] = src;
      float x[M::MAXS], z[NZ];
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = XA[s * (int64_t)J + src];
      M::draw(P, split_at(sk, (uint64_t)j), z);
      MP_UNROLL for (int i = 0; i < M::NOISE; ++i) zlog[((int64_t)t * M::NOISE + i) * J + j] = z[i];
      M::step_z(md, P, z, ts, dt, x);
      const float m = M::dmeas(md, P, y, x, tt);
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) XB[s * (int64_t)J + j] = x[s];
      lw[j] = m;
      lmax = nanmax(lmax, m);
    }
    lmax = block_nanmax(lmax, red);
    total = block_fixed_scan_seg(lw, Cc, J, lmax, tab);
    if (tid == 0) {
      stat[2 * t] = (double)lmax;
      stat[2 * t + 1] = ldexp((double)total, -sh);  // sum of exp(m - lmax)
    }
    if (tid < WAVE) loglik += step_loglik(lmax, total, sh) - (double)logJ;  // read by thread 0 only
    if (t < T - 1) ures = ures_slot[t & 1];
    float* tmp = XA;
    XA = XB;
    XB = tmp;
  }

  // ---------------- reverse sweep ----------------
  double gth[M::MAXP];
  MP_UNROLL for (int p = 0; p < M::MAXP; ++p) gth[p] = 0.0;
  for (int j = tid; j < J; j += B) {
    MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) xca[s * (int64_t)J + j] = 0.f;
    lama[j] = 0.f;
  }
  const int nblk = (T + K - 1) / K;
  for (int kb = nblk - 1; kb >= 0; --kb) {
    const int t0 = kb * K, n = min(T, t0 + K) - t0;
    const float* ck = ckx + (int64_t)kb * SJ;
    auto state = [&](int i) -> float* { return i == 0 ? const_cast<float*>(ck) : bx + (int64_t)(i - 1) * SJ; };
    // recompute the block's states from its checkpoint (the densities are
    // evaluated where the backward step needs them)
    for (int i = 0; i < n; ++i) {
      const int t = t0 + i;
      const float dt = dts[tidx0[t]];
      const float ts = tstart(t);
      const float* xin = state(i);
      float* xout = state(i + 1);
      const int32_t* anct = ancl + (int64_t)t * J;
      for (int j = tid; j < J; j += B) {
        const int src = anct[j];
        float x[M::MAXS], z[NZ];
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = xin[s * (int64_t)J + src];
        MP_UNROLL for (int q = 0; q < M::NOISE; ++q) z[q] = zlog[((int64_t)t * M::NOISE + q) * J + j];
        M::step_z(md, P, z, ts, dt, x);
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) xout[s * (int64_t)J + j] = x[s];
      }
      __syncthreads();
    }
    // backward through the block
    for (int i = n - 1; i >= 0; --i) {
      const int t = t0 + i;
      const Key sk = step_key(t);
      const float dt = dts[tidx0[t]];
      const float ts = tstart(t);
      const float* y = ys + (int64_t)t * md.dy;
      const float* xin = state(i);
      const float* xt = state(i + 1);
      const int32_t* anct = ancl + (int64_t)t * J;
      const float lmax = (float)stat[2 * t];
      const float inv_se = (float)(1.0 / stat[2 * t + 1]);
      for (int j = tid; j < J; j += B) {
        float x[M::MAXS], gx[M::MAXS], xb[M::MAXS], xf[M::MAXS], xfb[M::MAXS];
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = xt[s * (int64_t)J + j];
        // the forward's density, evaluated by the same operations
        const float e = M::dmeas_aux(P, x);
        const float mj = M::dmeas_a(md, P, y, x, e);
        const float pi = (lmax == -__builtin_inff() || mj != mj) ? 0.f : expf(mj - lmax) * inv_se;
        // this particle's carried adjoints are consumed here; as a parent it
        // starts the scatter below at zero
        const float lam = lama[j];
        float xc[M::MAXS];
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) {
          xc[s] = xca[s * (int64_t)J + j];
          xca[s * (int64_t)J + j] = 0.f;
        }
        lama[j] = 0.f;
        const float gbar = pi + lam;
        M::dmeas_grad(md, P, y, x, e, gbar, gx, gth);
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) xb[s] = gbar * gx[s] + xc[s];
        const int p = anct[j];
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) xf[s] = xin[s * (int64_t)J + p];
        M::step_vjp(md, P, split_at(sk, (uint64_t)j), xf, x, xb, xfb, gth);
        if (t == 0) {
          M::rinit_vjp(md, P, split_at(fkey, 1 + (uint64_t)j), a.ser.t0, xfb, gth);
        } else {
          MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) cx[s * (int64_t)J + j] = xfb[s];
          cl[j] = a.alpha * (pi - invJ + lam);
          ap[j] = p;
        }
      }
      (void)dt;
      (void)ts;
      if (t > 0) {
        // a_t is non-decreasing, so every parent's children are one run: the
        // first child of each run sums it in order and writes the parent's
        // adjoints (parents without children stay zero)
        __syncthreads();
        for (int c = tid; c < J; c += B) {
          const int q = ap[c];
          if (c > 0 && ap[c - 1] == q) continue;
          float sx[M::MAXS];
          MP_UNROLL for (int s = 0; s < M::MAXS; ++s) sx[s] = 0.f;
          float sl = 0.f;
          int k = c;
          do {
            MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) sx[s] += cx[s * (int64_t)J + k];
            sl += cl[k];
            ++k;
          } while (k < J && ap[k] == q);
          MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) xca[s * (int64_t)J + q] = sx[s];
          lama[q] = sl;
        }
      }
      __syncthreads();
    }
  }
  // d loglik / d theta_est
  block_sum_n<M::MAXP>(gth, red);
  if (tid == 0) {
    float jac[M::MAXP];
    if (md.identity_trans) {
      MP_UNROLL for (int p = 0; p < M::MAXP; ++p) jac[p] = 1.0f;
    } else {
      M::from_est_jac(md, th_est, jac);
    }
    MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < d)
      a.grad[(int64_t)f * d + p] = (float)(gth[p] * (double)jac[p]);
    a.loglik[f] = loglik;
  }
  __syncthreads();
}

// LDSW: the per-step working arrays live in LDS.  A compile-time choice, so
// their accesses compile to LDS instructions (a run-time choice between LDS
// and global memory would make every access a generic flat one).
template <class M, int NT, bool LDSW>
__device__ __forceinline__ void mop_grad_body(const GradArgs& a) {
  extern __shared__ __align__(256) char smem[];
  double* red = reinterpret_cast<double*>(smem);
  uint64_t* tab = reinterpret_cast<uint64_t*>(smem + kRedBytes);
  char* lwork = smem + kRedBytes + batched_tab_bytes(a.J, NT);
  for (int f = blockIdx.x; f < a.R; f += gridDim.x) {
    char* buf = a.ws + (int64_t)f * a.ws_stride;
    if constexpr (LDSW) {
      mop_grad_filter<M>(a, f, buf, lwork, red, tab);
    } else {
      const int S = M::FIXED_S ? M::MAXS : a.md.S;
      mop_grad_filter<M>(a, f, buf, buf + grad_scratch_layout(S, M::NOISE, a.J, a.ser.Tmax, a.K, false).work, red,
                         tab);
    }
  }
}

// 256 threads: the occupancy the registers allow (3 waves per SIMD).
template <class M, bool LDSW>
__global__ __launch_bounds__(256) void k_mop_grad(const GradArgs a) {
  mop_grad_body<M, 256, LDSW>(a);
}

// 512 threads capped at 128 VGPRs, so two workgroups (16 waves) share a CU:
// two filters per CU at 4 waves per SIMD.
template <class M, bool LDSW>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_mop_grad_wide(const GradArgs a) {
  mop_grad_body<M, 512, LDSW>(a);
}

template <class M, bool OK = M::DIFFERENTIABLE>
struct GradLauncher {
  static hipError_t run(const GradArgs&, int, int, size_t, hipStream_t) { return hipErrorNotSupported; }
};
template <class M>
struct GradLauncher<M, true> {
  static hipError_t run(const GradArgs& a, int block, int grid, size_t lds, hipStream_t st) {
    switch (block) {
      case 256:
        if (a.lds_work) mp_launch(k_mop_grad<M, true>, dim3(grid), dim3(256), lds, st, a);
        else mp_launch(k_mop_grad<M, false>, dim3(grid), dim3(256), lds, st, a);
        break;
      case 512:
        if (a.lds_work) mp_launch(k_mop_grad_wide<M, true>, dim3(grid), dim3(512), lds, st, a);
        else mp_launch(k_mop_grad_wide<M, false>, dim3(grid), dim3(512), lds, st, a);
        break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
};

// LDS bytes of a launch: reductions, scan table and (when resident) the
// per-step working arrays.
size_t mop_grad_lds_bytes(int S, int J, int block, bool lds_work) {
  return kRedBytes + batched_tab_bytes(J, block) + (lds_work ? (size_t)grad_work_layout(S, J).total : 0);
}

hipError_t launch_mop_grad(const GradArgs& a, int block, int grid, hipStream_t st) {
  const size_t lds = mop_grad_lds_bytes(a.md.S, a.J, block, a.lds_work);
  switch (a.md.kind) {
    case MODEL_SV: return GradLauncher<SV>::run(a, block, grid, lds, st);
    case MODEL_LG:
      if (a.md.dx == 1 && a.md.dy == 1) return GradLauncher<LG1>::run(a, block, grid, lds, st);
      if (a.md.dx == 2 && a.md.dy == 2) return GradLauncher<LG2>::run(a, block, grid, lds, st);
      // no adjoint above dx = dy = 2 (mop_grad_supported); not instantiated for 8
      if (a.md.dx > 4 || a.md.dy > 4) return hipErrorNotSupported;
      return GradLauncher<LG4>::run(a, block, grid, lds, st);
    case MODEL_SEIR: return GradLauncher<SEIR>::run(a, block, grid, lds, st);
  }
  return hipErrorInvalidValue;
}

bool mop_grad_supported(const ModelDesc& md) {
  switch (md.kind) {
    case MODEL_SV: return SV::DIFFERENTIABLE;
    case MODEL_LG:
      if (md.dx == 1 && md.dy == 1) return LG1::DIFFERENTIABLE;
      if (md.dx == 2 && md.dy == 2) return LG2::DIFFERENTIABLE;
      return false;
    case MODEL_SEIR: return SEIR::DIFFERENTIABLE;
  }
  return false;
}

int64_t mop_grad_scratch_bytes(int S, int nz, int J, int T, int K, bool lds_work) {
  return grad_scratch_layout(S, nz, J, T, K, lds_work).total;
}

// Noise words per particle-step the gradient logs for this model.
int mop_grad_noise_words(const ModelDesc& md) {
  switch (md.kind) {
    case MODEL_SV: return SV::NOISE;
    case MODEL_LG:
      if (md.dx == 1 && md.dy == 1) return LG1::NOISE;
      if (md.dx == 2 && md.dy == 2) return LG2::NOISE;
      return 0;
  }
  return 0;
}

}  // namespace mp
