// Stochastic SEIR with Euler-binomial transitions (the epidemic workload),
// written in the style of the reference's SIR model (pypomp/models/sir.py:171-249)
// and its binomial sampler (pypomp/random/binom.py:130-522).
//
// State words (float32 holding exact integer counts < 2^24): S, E, I, R, H,
// where H (weekly incidence, accumulator reset at each observation) counts
// E -> I transitions.  Parameters (natural scale, canonical order):
//   beta, amp, sigma, gamma, rho, k, N, S0f, E0, I0
// Estimation scale: log for beta, sigma, gamma, k, N, E0, I0; logit for amp, rho, S0f.
// Sub-step (rproc) with particle key `key`:
//   (kS, kE, kI) = split(key, 3)
//   beta_t = beta (1 + amp cos(2 pi t / 52)),   foi = beta_t I / N
//   dSE ~ Bin(S, 1 - exp(-foi dt)), dEI ~ Bin(E, 1 - exp(-sigma dt)),
//   dIR ~ Bin(I, 1 - exp(-gamma dt)), each drawn by fast_binomial from one
//   uniform of its key (binom.py:211-233).
// rinit: S = round(S0f N), E = round(E0), I = round(I0), R = N - S - E - I, H = 0.
// dmeas: y ~ NegBin(mean mu = max(rho H, 1e-10), size k)  (sir.py:236-249 form).
#pragma once

#include "models.hpp"

namespace mp {

// Normal quantile for u in (0, 1): sqrt(2) * erfinv(2u - 1) with Giles'
// single-precision erfinv, its argument w = -log((1-x)(1+x)) formed directly
// as -log(4 u (1 - u)) so no precision is lost near the tails.  The caller
// clips u to [eps, 1 - eps], so the log's argument is a positive normal number.
__device__ __forceinline__ float ndtri_f32(float u) {
  float w = -d_log_pos(4.0f * u * (1.0f - u));
  float x = 2.0f * u - 1.0f;
  float p;
  if (w < 5.0f) {
    w = w - 2.5f;
    p = 2.81022636e-08f;
    p = 3.43273939e-07f + p * w;
    p = -3.5233877e-06f + p * w;
    p = -4.39150654e-06f + p * w;
    p = 0.00021858087f + p * w;
    p = -0.00125372503f + p * w;
    p = -0.00417768164f + p * w;
    p = 0.246640727f + p * w;
    p = 1.50140941f + p * w;
  } else {
    w = sqrtf(w) - 3.0f;
    p = -0.000200214257f;
    p = 0.000100950558f + p * w;
    p = 0.00134934322f + p * w;
    p = -0.00367342844f + p * w;
    p = 0.00573950773f + p * w;
    p = -0.0076224613f + p * w;
    p = 0.00943887047f + p * w;
    p = 1.00167406f + p * w;
    p = 2.83297682f + p * w;
  }
  return 1.41421354f * (p * x);
}

// What the binomial inversion below needs of p alone: its first steps and the
// p-only factors of its correction terms and bottom-up stage.  BinomLazy forms
// each where the inversion uses it (no extra live registers); BinomP holds
// them, formed once, for draws that share p (a sub-step's E->I and I->R
// probabilities when every particle has the same parameters).  Both take the
// same operations, so the draws agree bit for bit.
struct BinomLazy {
  float p;  // the probability as given (edge cases)
  __device__ bool flip() const { return p > 0.5f; }
  __device__ float ps() const { return fminf(fmaxf(flip() ? 1.0f - p : p, 0.f), 1.f); }  // min(p, 1-p)
  __device__ float q() const { return 1.0f - ps(); }
  __device__ float pq() const { return ps() * q(); }
  __device__ float c6a() const { return 2.0f + 2.0f * ps(); }
  __device__ float c6b() const { return q() - ps(); }
  __device__ float c1a() const { return -2.0f + 14.0f * pq(); }
  __device__ float c1b() const { return -1.0f - 2.0f * pq(); }
  __device__ float c2() const { return (ps() - q()) * (2.0f + pq()); }
  __device__ float ratio() const { return ps() / fminf(fmaxf(1.0f - ps(), 1.17549435e-38f), 1.0f); }
  __device__ float logq() const { return d_log1p(-ps()); }
};

struct BinomP {
  float p_, ps_, q_, c6a_, c6b_, c1a_, c1b_, c2_, ratio_, logq_;
  bool flip_;
  __device__ float p() const { return p_; }
  __device__ bool flip() const { return flip_; }
  __device__ float ps() const { return ps_; }
  __device__ float q() const { return q_; }
  __device__ float c6a() const { return c6a_; }
  __device__ float c6b() const { return c6b_; }
  __device__ float c1a() const { return c1a_; }
  __device__ float c1b() const { return c1b_; }
  __device__ float c2() const { return c2_; }
  __device__ float ratio() const { return ratio_; }
  __device__ float logq() const { return logq_; }
};

__device__ __forceinline__ BinomP binom_p(float p) {
  const BinomLazy l{p};
  return BinomP{p, l.ps(), l.q(), l.c6a(), l.c6b(), l.c1a(), l.c1b(), l.c2(), l.ratio(), l.logq(), l.flip()};
}

__device__ __forceinline__ float prob(const BinomLazy& b) { return b.p; }
__device__ __forceinline__ float prob(const BinomP& b) { return b.p_; }

// Giles & Beentjes (2024) approximate inverse binomial CDF, order 2, with the
// exact bottom-up stage of at most 5 terms (binom.py:236-451), given p as a
// BinomLazy or BinomP.  Returns the count as a float; NaN for invalid arguments.
template <class B>
__device__ __forceinline__ float binominv_t(float u, float n, const B& b) {
  const float tiny = 1.17549435e-38f, eps = 1.1920929e-07f;
  const float p = prob(b);
  const bool invalid = (n < 0.f) || (p < 0.f) || (p > 1.f) || (u < 0.f) || (u > 1.f);
  const bool flip = b.flip();
  const float ps = b.ps();
  const float uf = flip ? 1.0f - u : u;
  const float ns = (n < 0.f) ? 1.0f : n;
  const float us = fminf(fmaxf(uf, eps), 1.0f - eps);
  const float q = b.q();
  const float w = ndtri_f32(us);
  const float w2 = w * w;
  const float np_ = ns * ps;
  const float npq = np_ * q;
  const float sq = sqrtf(fmaxf(npq, tiny));
  // the numerator is at least 2 (q >= ps): div_by is exact
  float q0 = np_ + sq * w + div_by(b.c6a() + b.c6b() * w2, 6.0f, 1.0f / 6.0f);
  float w3 = w2 * w;
  float q1 = q0 + (b.c1a() * w + b.c1b() * w3) / fmaxf(72.0f * sq, tiny);
  float w4 = w3 * w;
  float q2 = q1 + (b.c2() * (16.0f - 7.0f * w2 - 3.0f * w4)) / fmaxf(1620.0f * (sq * sq), tiny);
  float k = fminf(fmaxf(floorf(q2), 0.f), ns);
  k = fminf(k, np_ + 6.0f * sq + 5.0f);
  if (k < 10.0f || np_ < 4.0f) {
    // bottom-up exact CDF (binom.py:396-451)
    const float ue = fminf(fmaxf(uf, 0.f), 1.f);
    const float ratio = b.ratio();
    float pmf = (ns == 0.f) ? 1.0f : d_exp(ns * b.logq());
    float cdf = pmf;
    bool found = cdf >= ue;
    float res = found ? 0.f : k;
#pragma unroll
    for (int i = 1; i < 5; ++i) {
      float num = fmaxf(ns - (float)(i - 1), 0.f);
      // num is an integer count: num / 3 in Markstein's three operations equals
      // the IEEE quotient for every integer 0 .. 2^24 (checked exhaustively);
      // the divisions by 1, 2 and 4 are exact multiplications
      const float frac = i == 3 ? div_by(num, 3.0f, 1.0f / 3.0f) : num / (float)i;
      pmf = pmf * frac * ratio;
      cdf = cdf + pmf;
      bool stalled = (cdf > 1.0f - eps * 10.0f) && (pmf < eps);
      bool fnow = (cdf >= ue) || stalled;
      if (!found && fnow) res = (float)i;
      found = found || fnow;
    }
    k = fminf(fmaxf(res, 0.f), ns);
  }
  float r = flip ? ns - k : k;
  if (n == 0.f) r = 0.f;
  if (u == 0.f) r = 0.f;
  if (u == 1.f) r = ns;
  if (p == 0.f) r = 0.f;
  if (p == 1.f) r = ns;
  r = fminf(fmaxf(r, 0.f), ns);
  if (invalid) r = __builtin_nanf("");
  return r;
}

__device__ __forceinline__ float binominv_f32(float u, float n, float p) { return binominv_t(u, n, BinomLazy{p}); }

// fast_binomial(key, n, p) with a scalar draw (binom.py:211-233), int result,
// given p as a BinomLazy or BinomP.
template <class B>
__device__ __forceinline__ int fast_binomial_i(Key k, int n, const B& b) {
  float u = uniform01(bits_at(k, 0));
  u = fminf(fmaxf(u, 1.17549435e-38f), 0.99999994f);
  float r = binominv_t(u, (float)n, b);
  return r != r ? -1 : (int)r;
}

// NegBin(size k, mean mu) by inversion of u: the smallest y >= 0 with
// F(y) >= u, summing the pmf upward in float64 from pmf(0) = p^k with
// pmf(y+1) = pmf(y) (y + k)/(y + 1) q, p = k/(k+mu), q = mu/(k+mu).  pmf and F
// are carried in units of pmf(0) 2^(600 r) (r rescalings), so p^k may
// underflow.  Equals scipy's nbinom.ppf except where u lies within rounding of
// a CDF step; 0 at u = 0; NaN for k <= 0 or mu not finite.  O(y) steps, for
// data generation only.
__device__ inline double negbin_inverse(double u, double k, double mu) {
#pragma clang fp contract(off)
  if (!(k > 0.0) || !(mu >= 0.0) || mu == __builtin_inf()) return __builtin_nan("");
  if (!(u > 0.0)) return 0.0;
  const double p = k / (k + mu), q = mu / (k + mu);
  double lt = dd_log(u) - k * dd_log(p);
  double term = 1.0, cdf = 1.0, y = 0.0;
  double target = lt > 700.0 ? __builtin_inf() : dd_exp(lt);
  while (cdf < target) {
    const double ratio = (y + k) / (y + 1.0);
    term = (term * ratio) * q;
    cdf = cdf + term;
    y = y + 1.0;
    if (lt > 700.0 && cdf > 0x1p600) {
      term = term * 0x1p-600;
      cdf = cdf * 0x1p-600;
      lt = lt - 600 * 0.6931471805599453;
      if (lt <= 700.0) target = dd_exp(lt);
    }
    if ((term == 0.0 && y > mu) || y >= 1e9) break;
  }
  return y;
}

struct SEIR {
  static constexpr int MAXS = 5;
  static constexpr bool FIXED_S = true;
  static constexpr int MAXP = 10;
  static constexpr bool DIFFERENTIABLE = false;  // integer-valued transitions
  // the binomial draws depend on the state: no noise can be drawn ahead of it
  static constexpr int NOISE = 0;
  struct P {
    float beta, amp, sigma, gamma, rho, k, N, S0f, E0, I0;
    float invN;
    double lgk;
  };

  __host__ __device__ static float logit_inv(float x) { return 1.0f / (1.0f + d_exp(-x)); }

  __host__ __device__ static void from_est(const ModelDesc&, float* th) {
    th[0] = d_exp(th[0]);
    th[1] = logit_inv(th[1]);
    th[2] = d_exp(th[2]);
    th[3] = d_exp(th[3]);
    th[4] = logit_inv(th[4]);
    th[5] = d_exp(th[5]);
    th[6] = d_exp(th[6]);
    th[7] = logit_inv(th[7]);
    th[8] = d_exp(th[8]);
    th[9] = d_exp(th[9]);
  }

  __device__ static void prepare(const ModelDesc&, const float* th, P& p) {
    p.beta = th[0];
    p.amp = th[1];
    p.sigma = th[2];
    p.gamma = th[3];
    p.rho = th[4];
    p.k = th[5];
    p.N = th[6];
    p.S0f = th[7];
    p.E0 = th[8];
    p.I0 = th[9];
    p.invN = 1.0f / p.N;
    p.lgk = dd_lgamma((double)p.k);
  }

  __device__ static void rinit(const ModelDesc&, const P& p, Key, float, float* x) {
    float S = rintf(p.S0f * p.N), E = rintf(p.E0), I = rintf(p.I0);
    x[0] = S;
    x[1] = E;
    x[2] = I;
    x[3] = p.N - S - E - I;
    x[4] = 0.f;
  }

  __device__ static void draw(const P&, Key, float*) {}
  __device__ static void step_z(const ModelDesc&, const P&, const float*, float, float, float*) {}

  // Per-sub-step and per-observation values that do not depend on the
  // particle: a filter whose particles share one parameter set computes them
  // once per sub-step / observation (step_sub, dmeas_obs) with the same
  // operations as step / dmeas.  Tab and TabObs are the parts that depend on
  // the time or the data only, shared even when every particle carries its own
  // parameters (IF2).
  static constexpr bool HOIST = true;
  struct Tab {
    float cs;  // seasonal cosine at the sub-step's start time
  };
  struct TabObs {
    double lgy1;  // lgamma(y + 1)
  };
  struct Sub {
    float bt;
    BinomP bEI, bIR;  // binom_p of the E->I and I->R probabilities
  };
  struct Obs {
    double c;  // lgamma(y + k) - lgamma(k) - lgamma(y + 1)
  };

  __device__ static Tab tab_sub(float t) { return Tab{d_cos(6.28318531f * t / 52.0f)}; }

  __device__ static Sub substep_tab(const P& p, const Tab& tb, float dt) {
    Sub sb;
    sb.bt = p.beta * (1.0f + p.amp * tb.cs);
    sb.bEI = binom_p(-d_expm1(-p.sigma * dt));
    sb.bIR = binom_p(-d_expm1(-p.gamma * dt));
    return sb;
  }

  __device__ static Sub substep(const P& p, float t, float dt) { return substep_tab(p, tab_sub(t), dt); }

  // One Euler-binomial sub-step given the seasonal transmission rate bt and
  // the E->I and I->R probabilities (BinomP from the sub-step table, or
  // BinomLazy from the particle's own parameters).
  template <class B>
  __device__ static void step_core(const P& p, float bt, Key key, float dt, float* x, const B& bEI, const B& bIR) {
    int S = (int)x[0], E = (int)x[1], I = (int)x[2], R = (int)x[3], H = (int)x[4];
    const float foi = bt * (float)I * p.invN;
    const float pSE = -d_expm1(-foi * dt);
    const int dSE = fast_binomial_i(split_at(key, 0), S, BinomLazy{pSE});
    const int dEI = fast_binomial_i(split_at(key, 1), E, bEI);
    const int dIR = fast_binomial_i(split_at(key, 2), I, bIR);
    S -= dSE;
    E += dSE - dEI;
    I += dEI - dIR;
    R += dIR;
    H += dEI;
    x[0] = (float)S;
    x[1] = (float)E;
    x[2] = (float)I;
    x[3] = (float)R;
    x[4] = (float)H;
  }

  // the sub-step from the shared table (every particle has the parameters)
  __device__ static void step_sub(const P& p, const Sub& sb, Key key, float dt, float* x) {
    step_core(p, sb.bt, key, dt, x, sb.bEI, sb.bIR);
  }

  // the sub-step from its time-only table, with the particle's own parameters
  __device__ static void step_tab(const P& p, const Tab& tb, Key key, float dt, float* x) {
    const float bt = p.beta * (1.0f + p.amp * tb.cs);
    step_core(p, bt, key, dt, x, BinomLazy{-d_expm1(-p.sigma * dt)}, BinomLazy{-d_expm1(-p.gamma * dt)});
  }

  __device__ static void step(const ModelDesc&, const P& p, Key key, float t, float dt, float* x) {
    step_tab(p, tab_sub(t), key, dt, x);
  }

  __device__ static TabObs tab_obs(const float* y) { return TabObs{dd_lgamma((double)rintf(y[0]) + 1.0)}; }

  __device__ static Obs obs_tab(const P& p, const TabObs& to, const float* y) {
    const double yy = (double)rintf(y[0]);
    const double k = (double)p.k;
    return Obs{dd_lgamma(yy + k) - p.lgk - to.lgy1};
  }

  __device__ static Obs obs(const P& p, const float* y) { return obs_tab(p, tab_obs(y), y); }

  // NegBin log pmf evaluated in float64 (its three log-gamma terms cancel to
  // far below float32 resolution at counts ~1e4-1e5).
  __device__ static float dmeas_obs(const P& p, const Obs& o, const float* y, const float* x) {
    const double mu = (double)fmaxf(x[4] * p.rho, 1e-10f);
    const double yy = (double)rintf(y[0]);
    const double k = (double)p.k;
    const double v = o.c + k * dd_log(k / (k + mu)) + yy * dd_log(mu / (k + mu));
    return (float)v;
  }

  __device__ static float dmeas(const ModelDesc&, const P& p, const float* y, const float* x, float) {
    return dmeas_obs(p, obs(p, y), y, x);
  }

  // reports y ~ NegBin(mean max(rho H, 1e-10), size k) from one uniform
  __device__ static void rmeas(const ModelDesc&, const P& p, Key key, const float* x, float, float* y) {
    const double mu = fmax((double)(x[4] * p.rho), 1e-10);
    const double u = (double)uniform01(bits_at(key, 0));
    y[0] = (float)negbin_inverse(u, (double)p.k, mu);
  }
};

}  // namespace mp
