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
// Memory: the forward keeps only every K-th pre-resampling state and its
// ancestor indices; the reverse sweep recomputes each K-step block from its
// checkpoint (states, measurement densities and ancestry) and runs the block
// backwards.  Cost ~ 3 forward passes; storage O((T/K + K) J S).
#include "blockops.hpp"
#include "engine.hpp"
#include "resample.hpp"
#include "seir.hpp"

namespace mp {

struct GradScratch {
  int64_t ckx, cka, bx, bm, ba, c, lw, cx, cl, xca, lama, total;
};

__host__ __device__ inline GradScratch grad_scratch_layout(int S, int J, int T, int K) {
  GradScratch g;
  const int64_t nblk = (T + K - 1) / K;
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    int64_t r = o;
    o += align256(bytes);
    return r;
  };
  g.ckx = take(nblk * S * (int64_t)J * 4);
  g.cka = take(nblk * (int64_t)J * 4);
  g.bx = take((int64_t)(K + 1) * S * J * 4);
  g.bm = take((int64_t)K * J * 4);
  g.ba = take((int64_t)(K + 1) * J * 4);
  g.c = take((int64_t)J * 8);
  g.lw = take((int64_t)J * 4);
  g.cx = take((int64_t)S * J * 4);
  g.cl = take((int64_t)J * 4);
  g.xca = take((int64_t)S * J * 4);
  g.lama = take((int64_t)J * 4);
  g.total = o;
  return g;
}

// first index in [0, J) with a[i] >= v (a non-decreasing)
__device__ __forceinline__ int lower_bound_i32(const int32_t* a, int J, int v) {
  int lo = 0, hi = J;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <class M>
__device__ void mop_grad_filter(const GradArgs& a, const int f, char* buf, double* red, uint64_t* tab) {
  const ModelDesc md = a.md;
  const int J = a.J, S = M::FIXED_S ? M::MAXS : md.S, d = md.d, T = a.ser.Tmax, K = a.K;
  const int tid = threadIdx.x, B = blockDim.x;
  const GradScratch L = grad_scratch_layout(S, J, T, K);
  float* ckx = reinterpret_cast<float*>(buf + L.ckx);
  int32_t* cka = reinterpret_cast<int32_t*>(buf + L.cka);
  float* bx = reinterpret_cast<float*>(buf + L.bx);
  float* bm = reinterpret_cast<float*>(buf + L.bm);
  int32_t* ba = reinterpret_cast<int32_t*>(buf + L.ba);
  uint64_t* Cc = reinterpret_cast<uint64_t*>(buf + L.c);
  float* lw = reinterpret_cast<float*>(buf + L.lw);
  float* cx = reinterpret_cast<float*>(buf + L.cx);
  float* cl = reinterpret_cast<float*>(buf + L.cl);
  float* xca = reinterpret_cast<float*>(buf + L.xca);
  float* lama = reinterpret_cast<float*>(buf + L.lama);
  const int64_t SJ = (int64_t)S * J;
  const float* ys = a.ser.ys;
  const float* times = a.ser.times;
  const int32_t* tidx0 = a.ser.tidx0;
  const float* dts = a.ser.dt_ext;
  const uint32_t* sch = a.sched + 2 * (int64_t)f * a.sched_len;
  const Key fkey{a.keys[2 * f], a.keys[2 * f + 1]};
  const float fJ = (float)J, logJ = d_log(fJ), invJ = 1.0f / fJ;

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
    return Key{sch[2 * base], sch[2 * base + 1]};
  };
  auto res_key = [&](int t) -> Key {
    const int64_t base = tidx0[t] + (int64_t)t + 1;  // one sub-step per observation
    return Key{sch[2 * base], sch[2 * base + 1]};
  };
  auto tstart = [&](int t) { return t == 0 ? a.ser.t0 : times[t - 1]; };

  // ---------------- forward: value + checkpoints ----------------
  float* XA = bx;
  float* XB = bx + SJ;
  for (int j = tid; j < J; j += B) {
    float x[M::MAXS];
    M::rinit(md, P, split_at(fkey, 1 + (uint64_t)j), a.ser.t0, x);
    MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) XA[s * (int64_t)J + j] = x[s];
  }
  __syncthreads();
  double loglik = 0.0;
  uint64_t total = 0;
  float ures = 0.f;
  for (int t = 0; t < T; ++t) {
    const bool ck = (t % K) == 0;
    const int kb = t / K;
    const Key sk = step_key(t);
    const float dt = dts[tidx0[t]];
    const float* y = ys + (int64_t)t * md.dy;
    float lmax = -__builtin_inff(), lmin = __builtin_inff();
    for (int j = tid; j < J; j += B) {
      const int src = t == 0 ? j : ancestor_of(Cc, J, ures, total, j);
      if (ck) {
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S)
          ckx[(int64_t)kb * SJ + s * (int64_t)J + j] = XA[s * (int64_t)J + j];
        cka[(int64_t)kb * J + j] = src;
      }
      float x[M::MAXS];
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = XA[s * (int64_t)J + src];
      M::step(md, P, split_at(sk, (uint64_t)j), tstart(t), dt, x);
      const float m = M::dmeas(md, P, y, x, times[t]);
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) XB[s * (int64_t)J + j] = x[s];
      lw[j] = m;
      lmax = nanmax(lmax, m);
      lmin = nanmin(lmin, m);
    }
    block_maxmin(lmax, lmin, red);
    double se = 0.0;
    if (lmax == lmax && lmax != -__builtin_inff())
      for (int j = tid; j < J; j += B) {
        const float v = lw[j];
        if (v == v) se += (double)expf(v - lmax);
      }
    se = block_sum(se, red);
    loglik += (lmax == -__builtin_inff()) ? -__builtin_inf() : ((double)lmax + log(se) - (double)logJ);
    if (t < T - 1) {
      total = block_fixed_scan(lw, Cc, J, lmax, tab);
      ures = uniform01(bits_at(res_key(t), 0));
    } else {
      __syncthreads();
    }
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
    const int t0 = kb * K, t1 = min(T, t0 + K), n = t1 - t0;
    // recompute the block from its checkpoint
    for (int j = tid; j < J; j += B) {
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S)
        bx[s * (int64_t)J + j] = ckx[(int64_t)kb * SJ + s * (int64_t)J + j];
      ba[j] = cka[(int64_t)kb * J + j];
    }
    __syncthreads();
    for (int i = 0; i < n; ++i) {
      const int t = t0 + i;
      const Key sk = step_key(t);
      const float dt = dts[tidx0[t]];
      const float* y = ys + (int64_t)t * md.dy;
      const float* xin = bx + (int64_t)i * SJ;
      float* xout = bx + (int64_t)(i + 1) * SJ;
      const int32_t* anc = ba + (int64_t)i * J;
      float lmax = -__builtin_inff(), lmin = __builtin_inff();
      for (int j = tid; j < J; j += B) {
        const int src = anc[j];
        float x[M::MAXS];
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = xin[s * (int64_t)J + src];
        M::step(md, P, split_at(sk, (uint64_t)j), tstart(t), dt, x);
        const float m = M::dmeas(md, P, y, x, times[t]);
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) xout[s * (int64_t)J + j] = x[s];
        bm[(int64_t)i * J + j] = m;
        lw[j] = m;
        lmax = nanmax(lmax, m);
        lmin = nanmin(lmin, m);
      }
      block_maxmin(lmax, lmin, red);
      if (i < n - 1) {
        const uint64_t tot = block_fixed_scan(lw, Cc, J, lmax, tab);
        const float u = uniform01(bits_at(res_key(t), 0));
        int32_t* nanc = ba + (int64_t)(i + 1) * J;
        for (int j = tid; j < J; j += B) nanc[j] = ancestor_of(Cc, J, u, tot, j);
      }
      __syncthreads();
    }
    // backward through the block
    for (int i = n - 1; i >= 0; --i) {
      const int t = t0 + i;
      const Key sk = step_key(t);
      const float dt = dts[tidx0[t]];
      const float* y = ys + (int64_t)t * md.dy;
      const float* xin = bx + (int64_t)i * SJ;
      const float* xt = bx + (int64_t)(i + 1) * SJ;
      const float* mt = bm + (int64_t)i * J;
      const int32_t* anc = ba + (int64_t)i * J;
      float lmax = -__builtin_inff(), lmin = __builtin_inff();
      for (int j = tid; j < J; j += B) lmax = nanmax(lmax, mt[j]);
      block_maxmin(lmax, lmin, red);
      double se = 0.0;
      if (lmax == lmax && lmax != -__builtin_inff())
        for (int j = tid; j < J; j += B) {
          const float v = mt[j];
          if (v == v) se += (double)expf(v - lmax);
        }
      se = block_sum(se, red);
      const float inv_se = (float)(1.0 / se);
      for (int j = tid; j < J; j += B) {
        const float mj = mt[j];
        const float pi = (lmax == -__builtin_inff() || mj != mj) ? 0.f : expf(mj - lmax) * inv_se;
        const float lam = lama[j];
        const float gbar = pi + lam;
        float x[M::MAXS], gx[M::MAXS], xb[M::MAXS], xf[M::MAXS], xfb[M::MAXS];
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = xt[s * (int64_t)J + j];
        M::dmeas_grad(md, P, y, x, gbar, gx, gth);
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) xb[s] = gbar * gx[s] + xca[s * (int64_t)J + j];
        const int p = anc[j];
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) xf[s] = xin[s * (int64_t)J + p];
        M::step_vjp(md, P, split_at(sk, (uint64_t)j), tstart(t), dt, xf, xb, xfb, gth);
        if (t == 0) {
          M::rinit_vjp(md, P, split_at(fkey, 1 + (uint64_t)j), a.ser.t0, xfb, gth);
        } else {
          MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) cx[s * (int64_t)J + j] = xfb[s];
          cl[j] = a.alpha * (pi - invJ + lam);
        }
      }
      __syncthreads();
      if (t > 0) {
        // parents q collect their children's contributions (contiguous runs of anc)
        for (int q = tid; q < J; q += B) {
          const int lo = lower_bound_i32(anc, J, q);
          const int hi = lower_bound_i32(anc, J, q + 1);
          float sx[M::MAXS];
          MP_UNROLL for (int s = 0; s < M::MAXS; ++s) sx[s] = 0.f;
          float sl = 0.f;
          for (int c = lo; c < hi; ++c) {
            MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) sx[s] += cx[s * (int64_t)J + c];
            sl += cl[c];
          }
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

template <class M>
__global__ __launch_bounds__(256) void k_mop_grad(const GradArgs a) {
  extern __shared__ __align__(256) char smem[];
  double* red = reinterpret_cast<double*>(smem);
  uint64_t* tab = reinterpret_cast<uint64_t*>(smem + kRedBytes);
  for (int f = blockIdx.x; f < a.R; f += gridDim.x) mop_grad_filter<M>(a, f, a.ws + (int64_t)f * a.ws_stride, red, tab);
}

template <class M, bool OK = M::DIFFERENTIABLE>
struct GradLauncher {
  static hipError_t run(const GradArgs&, int, int, size_t, hipStream_t) { return hipErrorNotSupported; }
};
template <class M>
struct GradLauncher<M, true> {
  static hipError_t run(const GradArgs& a, int block, int grid, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL(k_mop_grad<M>, dim3(grid), dim3(block), lds, st, a);
    return hipGetLastError();
  }
};

hipError_t launch_mop_grad(const GradArgs& a, int block, int grid, hipStream_t st) {
  const size_t lds = kRedBytes + batched_tab_bytes(a.J, block);
  switch (a.md.kind) {
    case MODEL_SV: return GradLauncher<SV>::run(a, block, grid, lds, st);
    case MODEL_LG:
      if (a.md.dx == 1 && a.md.dy == 1) return GradLauncher<LG1>::run(a, block, grid, lds, st);
      if (a.md.dx == 2 && a.md.dy == 2) return GradLauncher<LG2>::run(a, block, grid, lds, st);
      return GradLauncher<LG4>::run(a, block, grid, lds, st);
    case MODEL_SEIR: return GradLauncher<SEIR>::run(a, block, grid, lds, st);
  }
  return hipErrorInvalidValue;
}

int64_t mop_grad_scratch_bytes(int S, int J, int T, int K) { return grad_scratch_layout(S, J, T, K).total; }

}  // namespace mp
