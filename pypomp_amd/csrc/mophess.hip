// Hessian of the MOP(alpha) log-likelihood with respect to the
// estimation-scale parameters: what jax.hessian computes through
// _mop_internal_mean (pypomp/core/algorithms/ad_helpers.py:8, 56-75;
// mop.py:19-158), ancestry and random numbers held fixed.
//
// jax.hessian is forward over reverse, and the tangent of stop_gradient is
// zero in both passes, so the Hessian is the second derivative of
//   F(theta; s) = sum_t LSE_j(wP_t,j + m_t,j) - LSE_j(wP_t,j),
//   wP_t = alpha wF_{t-1},   wF_t = (wP_t + m_t - s_t)[a_t]
// with s held at the densities' values at theta.  This kernel evaluates F in
// forward mode over hyper-dual numbers: one workgroup per (parameter row,
// parameter pair p <= q) seeds d/dtheta_p and d/dtheta_q and carries every
// particle's state and weight with their first and mixed second derivatives.
// Value parts are the primal's float32 operations in the primal's order (the
// same resampling ancestry as mp_mop); derivative parts are float64.
//
//   LSE derivatives, a_j hyper-dual, pi = softmax(a.v):
//     d_p LSE = sum pi a_j.p,   d_pq LSE = sum pi (a_j.pq + a_j.p a_j.q) - d_p LSE d_q LSE.
#include "blockops.hpp"
#include "engine.hpp"
#include "mech.hpp"
#include "resample.hpp"

namespace mp {

// Hyper-dual number: value v (float32, the primal's arithmetic) and the parts
// along the two seeded directions and their mixed second derivative (float64).
struct HD {
  float v;
  double a, b, ab;
};

__device__ __forceinline__ HD hd(float v) { return HD{v, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ HD operator+(const HD& x, const HD& y) {
  return HD{x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab};
}
__device__ __forceinline__ HD operator-(const HD& x, const HD& y) {
  return HD{x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab};
}
__device__ __forceinline__ HD operator*(const HD& x, const HD& y) {
  const double xv = x.v, yv = y.v;
  return HD{x.v * y.v, x.a * yv + xv * y.a, x.b * yv + xv * y.b, x.ab * yv + x.a * y.b + x.b * y.a + xv * y.ab};
}
__device__ __forceinline__ HD operator/(const HD& x, const HD& y) {
  const double yv = y.v, q = (double)x.v / yv;
  const double qa = (x.a - q * y.a) / yv, qb = (x.b - q * y.b) / yv;
  return HD{x.v / y.v, qa, qb, (x.ab - qa * y.b - qb * y.a - q * y.ab) / yv};
}
__device__ __forceinline__ HD operator+(const HD& x, float c) { return x + hd(c); }
__device__ __forceinline__ HD operator+(float c, const HD& x) { return hd(c) + x; }
__device__ __forceinline__ HD operator-(const HD& x, float c) { return x - hd(c); }
__device__ __forceinline__ HD operator-(float c, const HD& x) { return hd(c) - x; }
__device__ __forceinline__ HD operator*(float c, const HD& x) { return hd(c) * x; }
__device__ __forceinline__ HD operator*(const HD& x, float c) { return x * hd(c); }
__device__ __forceinline__ HD operator/(float c, const HD& x) { return hd(c) / x; }

// g(x) with g' and g'' at x.v: the chain rule to second order.
__device__ __forceinline__ HD hd_chain(float gv, const HD& x, double d1, double d2) {
  return HD{gv, d1 * x.a, d1 * x.b, d1 * x.ab + d2 * x.a * x.b};
}
__device__ __forceinline__ HD hd_exp(const HD& x) {
  const double e = exp((double)x.v);
  return hd_chain(d_exp(x.v), x, e, e);
}
__device__ __forceinline__ HD hd_log(const HD& x) {
  const double r = 1.0 / (double)x.v;
  return hd_chain(d_log(x.v), x, r, -r * r);
}
__device__ __forceinline__ HD hd_tanh(const HD& x) {
  const double t = tanh((double)x.v), d = 1.0 - t * t;
  return hd_chain(d_tanh(x.v), x, d, -2.0 * t * d);
}
__device__ __forceinline__ HD hd_sqrt(const HD& x) {
  const double s = sqrt((double)x.v);
  return hd_chain(sqrtf(x.v), x, 0.5 / s, -0.25 / (s * (double)x.v));
}

// ---- hyper-dual forms of the differentiable models -------------------------
// Each mirrors the primal functor's float operations (models.hpp) on HD.

// Stochastic volatility with leverage (SV in models.hpp).
struct SVH {
  static constexpr int S = 2, MAXP = 6, NZ = 1;
  struct PH {
    HD mu, phi, sigma, rho, V0, S0, c0, inv1mr2;
  };
  __device__ static void from_est(const ModelDesc&, HD* th) {
    th[1] = hd_tanh(th[1]);
    th[2] = hd_exp(th[2]);
    th[3] = hd_tanh(th[3]);
    th[5] = hd_exp(th[5]);
  }
  __device__ static void prepare(const ModelDesc&, const HD* th, PH& p) {
    p.mu = th[0];
    p.phi = th[1];
    p.sigma = th[2];
    p.rho = th[3];
    p.V0 = th[4];
    p.S0 = th[5];
    const HD omr2 = 1.0f - p.rho * p.rho;
    p.inv1mr2 = 1.0f / omr2;
    p.c0 = -0.918938533f - 0.5f * hd_log(omr2);
  }
  __device__ static void rinit(const ModelDesc&, const PH& p, Key k, HD* x) {
    const float z = normal_at(k, 0);
    x[0] = p.V0 + p.S0 * z;
    x[1] = hd(0.0f);
  }
  __device__ static void draw(const ModelDesc&, Key k, float* z) { z[0] = normal_at(k, 0); }
  __device__ static void step_z(const ModelDesc&, const PH& p, const float* z, HD* x) {
    x[0] = p.mu + p.phi * (x[0] - p.mu) + p.sigma * z[0];
    x[1] = hd(z[0]);
  }
  __device__ static HD dmeas(const ModelDesc&, const PH& p, const float* y, const HD* x) {
    const HD e = hd_exp(-0.5f * x[0]);
    const HD u = y[0] * e - p.rho * x[1];
    return p.c0 - 0.5f * x[0] - 0.5f * u * u * p.inv1mr2;
  }
};

// Linear Gaussian with dx = dy = MD (LGT<MD, true> in models.hpp).
template <int MD>
struct LGH {
  static constexpr int S = MD, MAXP = 2 * MD * MD + 3 * (MD * (MD + 1) / 2) + MD, NZ = MD;
  struct PH {
    HD A[MD][MD], C[MD][MD], LQ[MD][MD], L0[MD][MD], LC[MD][MD];
    HD X0[MD];
    HD logdet;
    float half_n_log2pi;
  };
  __device__ static void from_est(const ModelDesc&, HD* th) {
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      const int o = MD * MD + MD * MD + i * (i + 1) / 2 + i;
      th[o] = hd_exp(th[o]);
    }
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      const int o = MD * MD + MD * MD + MD * (MD + 1) / 2 + i * (i + 1) / 2 + i;
      th[o] = hd_exp(th[o]);
    }
  }
  __device__ static void prepare(const ModelDesc& md, const HD* th, PH& p) {
    constexpr int oC = MD * MD, oQ = oC + MD * MD, oR = oQ + MD * (MD + 1) / 2, oX = oR + MD * (MD + 1) / 2,
                  oP = oX + MD;
    HD LR[MD][MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) MP_UNROLL for (int j = 0; j < MD; ++j) {
      p.A[i][j] = th[i * MD + j];
      p.C[i][j] = th[oC + i * MD + j];
      p.LQ[i][j] = j <= i ? th[oQ + i * (i + 1) / 2 + j] : hd(0.0f);
      LR[i][j] = j <= i ? th[oR + i * (i + 1) / 2 + j] : hd(0.0f);
      p.LC[i][j] = hd(0.0f);
    }
    MP_UNROLL for (int i = 0; i < MD; ++i) MP_UNROLL for (int j = 0; j < MD; ++j)
      p.L0[i][j] = md.lg_sep_init ? (j <= i ? th[oP + i * (i + 1) / 2 + j] : hd(0.0f)) : p.LQ[i][j];
    MP_UNROLL for (int i = 0; i < MD; ++i) p.X0[i] = th[oX + i];
    HD cov[MD][MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) MP_UNROLL for (int j = 0; j < MD; ++j) {
      HD s = hd(0.0f);
      MP_UNROLL for (int k = 0; k < MD; ++k) s = s + LR[i][k] * LR[j][k];
      cov[i][j] = s;
    }
    HD logdet = hd(0.0f);
    MP_UNROLL for (int j = 0; j < MD; ++j) {
      HD s = cov[j][j];
      MP_UNROLL for (int k = 0; k < MD; ++k) if (k < j) s = s - p.LC[j][k] * p.LC[j][k];
      const HD djj = hd_sqrt(s);
      p.LC[j][j] = djj;
      logdet = logdet + hd_log(djj);
      MP_UNROLL for (int i = 0; i < MD; ++i) if (i > j) {
        HD t = cov[i][j];
        MP_UNROLL for (int k = 0; k < MD; ++k) if (k < j) t = t - p.LC[i][k] * p.LC[j][k];
        p.LC[i][j] = t / djj;
      }
    }
    p.logdet = logdet;
    p.half_n_log2pi = 0.5f * (float)MD * 1.83787707f;
  }
  __device__ static void rinit(const ModelDesc&, const PH& p, Key k, HD* x) {
    float z[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) z[i] = normal_at(k, i);
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      HD s = p.L0[i][0] * z[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) s = s + p.L0[i][j] * z[j];
      x[i] = p.X0[i] + s;
    }
  }
  __device__ static void draw(const ModelDesc&, Key k, float* z) {
    MP_UNROLL for (int i = 0; i < MD; ++i) z[i] = normal_at(k, i);
  }
  __device__ static void step_z(const ModelDesc&, const PH& p, const float* z, HD* x) {
    HD xn[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      HD a = p.A[i][0] * x[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) a = a + p.A[i][j] * x[j];
      HD s = p.LQ[i][0] * z[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) s = s + p.LQ[i][j] * z[j];
      xn[i] = a + s;
    }
    MP_UNROLL for (int i = 0; i < MD; ++i) x[i] = xn[i];
  }
  __device__ static HD dmeas(const ModelDesc&, const PH& p, const float* y, const HD* x) {
    HD r[MD], s[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      HD m = p.C[i][0] * x[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) m = m + p.C[i][j] * x[j];
      r[i] = y[i] - m;
    }
    HD quad = hd(0.0f);
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      HD a = r[i];
      MP_UNROLL for (int k = 0; k < MD; ++k) if (k < i) a = a - p.LC[i][k] * s[k];
      s[i] = a / p.LC[i][i];
      quad = quad + s[i] * s[i];
    }
    return -0.5f * quad - p.half_n_log2pi - p.logdet;
  }
};

__device__ __forceinline__ Key hs_key(const uint32_t* p, int64_t i) { return Key{p[2 * i], p[2 * i + 1]}; }

// Parameter pair pi of the upper triangle in row-major order: (p, q), p <= q.
__device__ __forceinline__ void pair_of(int pi, int d, int& p, int& q) {
  p = 0;
  int rem = pi;
  while (rem >= d - p) {
    rem -= d - p;
    ++p;
  }
  q = p + rem;
}

template <class H>
__device__ void run_hess_task(const HessArgs& a, int r, int pi, char* buf, double* red, uint64_t* tab,
                              typename H::PH& P) {
  const ModelDesc md = a.md;
  constexpr int S = H::S;
  const int J = a.J, d = md.d, dy = md.dy, tid = threadIdx.x, B = blockDim.x;
  const int T = a.ser.Tmax;
  const float* ys = a.ser.ys;
  const int32_t* nstep = a.ser.nstep;
  const int32_t* tidx0 = a.ser.tidx0;
  const uint32_t* sch = a.sched + 2 * (int64_t)r * a.sched_len;
  int pp, qq;
  pair_of(pi, d, pp, qq);

  const HessScratch L = hess_scratch_layout(S, J);
  HD* X[2] = {reinterpret_cast<HD*>(buf + L.off_x0), reinterpret_cast<HD*>(buf + L.off_x1)};
  HD* W[2] = {reinterpret_cast<HD*>(buf + L.off_w0), reinterpret_cast<HD*>(buf + L.off_w1)};
  float* lw = reinterpret_cast<float*>(buf + L.off_lw);
  uint64_t* Cc = reinterpret_cast<uint64_t*>(buf + L.off_c);

  if (tid == 0) {
    HD th[H::MAXP];
    MP_UNROLL for (int i = 0; i < H::MAXP; ++i) if (i < d) {
      th[i] = hd(a.theta[(int64_t)r * a.theta_stride + i]);
      th[i].a = i == pp ? 1.0 : 0.0;
      th[i].b = i == qq ? 1.0 : 0.0;
    }
    if (!md.identity_trans) H::from_est(md, th);
    H::prepare(md, th, P);
  }
  __syncthreads();

  const float fJ = (float)J;
  const Key fkey = hs_key(a.keys, r);
  const float w0 = d_log(1.0f / fJ);
  for (int j = tid; j < J; j += B) {
    HD x[S];
    H::rinit(md, P, split_at(fkey, 1 + (uint64_t)j), x);
    MP_UNROLL for (int s = 0; s < S; ++s) X[0][(int64_t)s * J + j] = x[s];
    W[0][j] = hd(w0);
  }
  __syncthreads();

  int cur = 0;
  bool have_anc = false;
  float ures = 0.0f;
  uint64_t total = 0;
  double loglik = 0.0, gp = 0.0, hpq = 0.0;
  const float logJ = d_log(fJ);
  const int sh = shift_for(J);
  for (int t = 0; t < T; ++t) {
    const int nst = nstep[t], ti0 = tidx0[t];
    const int64_t base = ti0 + (int64_t)t;
    const float* y = ys + (int64_t)t * dy;
    float lmax = -__builtin_inff();
    // LSE(wP): wP.v is one value for all particles, so its weights are 1/J
    double w2a = 0.0, w2b = 0.0, w2ab = 0.0;
    for (int j = tid; j < J; j += B) {
      const int src = have_anc ? ancestor_seg(Cc, tab, J, ures, total, j) : j;
      HD x[S];
      MP_UNROLL for (int s = 0; s < S; ++s) x[s] = X[cur][(int64_t)s * J + src];
      const HD wP = a.alpha * W[cur][src];
      for (int k = 0; k < nst; ++k) {
        float z[H::NZ];
        H::draw(md, split_at(hs_key(sch, base + k), (uint64_t)j), z);
        H::step_z(md, P, z, x);
      }
      const HD m = H::dmeas(md, P, y, x);
      MP_UNROLL for (int s = 0; s < S; ++s) X[cur ^ 1][(int64_t)s * J + j] = x[s];
      // wF before the gather: (wP + m - sg(m)), value wP.v
      W[cur ^ 1][j] = HD{wP.v, wP.a + m.a, wP.b + m.b, wP.ab + m.ab};
      lw[j] = m.v;
      lmax = nanmax(lmax, m.v);
      w2a += wP.a;
      w2b += wP.b;
      w2ab += wP.ab + wP.a * wP.b;
    }
    lmax = block_nanmax(lmax, red);
    // the value: the primal's exact fixed-point total (the same bits as mp_mop),
    // and the prefix sums the resampler searches
    total = block_fixed_scan_seg(lw, Cc, J, lmax, tab);
    loglik += step_loglik(lmax, total, sh) - (double)logJ;
    // LSE(wP + m): weights exp(m.v - max), as wP.v is common
    double acc[7] = {w2a, w2b, w2ab, 0.0, 0.0, 0.0, 0.0};
    for (int j = tid; j < J; j += B) {
      const double e = exp((double)lw[j] - (double)lmax);
      const HD wf = W[cur ^ 1][j];
      acc[3] += e;
      acc[4] += e * wf.a;
      acc[5] += e * wf.b;
      acc[6] += e * (wf.ab + wf.a * wf.b);
    }
    block_sum_n<7>(acc, red);
    const double iJ = 1.0 / (double)J, iW = 1.0 / acc[3];
    const double l1a = acc[4] * iW, l1ab = acc[6] * iW - l1a * (acc[5] * iW);
    const double l2a = acc[0] * iJ, l2ab = acc[2] * iJ - l2a * (acc[1] * iJ);
    gp += l1a - l2a;
    hpq += l1ab - l2ab;
    // resample every step (mop.py:122-123) from normalize(sg(m))
    ures = uniform01(bits_at(hs_key(sch, base + nst), 0));
    have_anc = true;
    cur ^= 1;
  }
  if (tid == 0) {
    if (pi == 0) a.loglik[r] = loglik;
    if (pp == qq) a.grad[(int64_t)r * d + pp] = (float)gp;
    a.hess[((int64_t)r * d + pp) * d + qq] = (float)hpq;
    a.hess[((int64_t)r * d + qq) * d + pp] = (float)hpq;
  }
  __syncthreads();
}

template <class H>
__global__ __launch_bounds__(256) void k_mop_hess(const HessArgs a) {
  extern __shared__ __align__(256) char smem[];
  __shared__ typename H::PH s_P;
  double* red = reinterpret_cast<double*>(smem);
  uint64_t* tab = reinterpret_cast<uint64_t*>(smem + kRedBytes);
  const int npairs = a.md.d * (a.md.d + 1) / 2;
  const int64_t ntasks = (int64_t)a.R * npairs;
  for (int64_t q = blockIdx.x; q < ntasks; q += gridDim.x)
    run_hess_task<H>(a, (int)(q / npairs), (int)(q % npairs), a.ws + (int64_t)blockIdx.x * a.ws_stride, red, tab,
                     s_P);
}

template <class H>
static hipError_t launch_hess_t(const HessArgs& a, int grid, hipStream_t st) {
  const size_t lds = kRedBytes + batched_tab_bytes(a.J, 256);
  mp_launch(k_mop_hess<H>, dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError();
}

hipError_t launch_mop_hess(const HessArgs& a, int grid, hipStream_t st) {
  switch (a.md.kind) {
    case MODEL_SV: return launch_hess_t<SVH>(a, grid, st);
    case MODEL_LG:
      if (a.md.dx == 1 && a.md.dy == 1) return launch_hess_t<LGH<1>>(a, grid, st);
      if (a.md.dx == 2 && a.md.dy == 2) return launch_hess_t<LGH<2>>(a, grid, st);
      return hipErrorNotSupported;
  }
  return hipErrorNotSupported;
}

bool mop_hess_supported(const ModelDesc& md) {
  return md.kind == MODEL_SV || (md.kind == MODEL_LG && md.dx == md.dy && (md.dx == 1 || md.dx == 2));
}

size_t mop_hess_lds_bytes(int J) { return kRedBytes + batched_tab_bytes(J, 256); }

}  // namespace mp
