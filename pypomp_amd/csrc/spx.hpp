// S&P 500 volatility model (pypomp/models/spx.py:56-151), native form.
//
// State (V, S); theta = (mu, kappa, theta, xi, rho, V_0); one covariate,
// y_prev, the previous observation's log return.  Per Euler sub-step, with
// dZ a standard normal:
//   dWs = (y_prev - mu + 0.5 V) / sqrt(V)
//   dWv = rho dWs + sqrt(1 - rho^2) dZ
//   S   = S + S (mu + sqrt(max(V, 1e-32)) dWs)
//   V   = max(V + kappa (theta - V) + xi sqrt(V) dWv, 1e-32)
// (spx.py:66-126), each operation in the reference's order, float32.
// rinit: V = V_0, S = 1105 (spx.py:56-63).  dmeas: log N(y; mu - V/2, sqrt(V))
// in jax.scipy.stats.norm.logpdf's arithmetic (log(2 pi s^2) + (y - loc)^2 /
// s^2) / -2 with s = sqrt(V) (spx.py:129-137).  The reference defines no rmeas;
// rmeas here draws from that density, y = (mu - V/2) + sqrt(V) e.
// Estimation scale (spx.py:140-151): log of mu, kappa, theta, xi and V_0;
// rho = -1 + 2 / (1 + exp(-x)).
//
// Draws: dZ is normal(key) of the particle key split(sk, J)[j] (the
// per-particle rproc, spx.py:99-126), or normal(sk, (J,))[j] of the sub-step
// key itself (the @vectorized rproc, spx.py:66-96).  The kernels reach the
// model through mech.hpp, which hands it the key and the draw index.
#pragma once

#include "models.hpp"

namespace mp {

struct SPX {
  static constexpr int MAXS = 2;
  static constexpr bool FIXED_S = true;
  static constexpr int MAXP = 6;
  static constexpr bool DIFFERENTIABLE = false;
  static constexpr bool HOIST = false;
  static constexpr int NOISE = 1;    // dZ depends on the key only
  static constexpr bool DRAW_USES_P = false;
  static constexpr bool EXT = true;  // reads covariates; defines the @vectorized draw
  static constexpr int DC = 1;       // covariates read: y_prev
  struct P {
    float mu, kappa, theta, xi, rho, V0;
    float sq1mr2;  // sqrt(1 - rho^2)
  };

  __host__ __device__ static void from_est(const ModelDesc&, float* th) {
    th[0] = d_exp(th[0]);
    th[1] = d_exp(th[1]);
    th[2] = d_exp(th[2]);
    th[3] = d_exp(th[3]);
    th[4] = -1.0f + 2.0f / (1.0f + d_exp(-th[4]));
    th[5] = d_exp(th[5]);
  }

  __device__ static void prepare(const ModelDesc&, const float* th, P& p) {
    p.mu = th[0];
    p.kappa = th[1];
    p.theta = th[2];
    p.xi = th[3];
    p.rho = th[4];
    p.V0 = th[5];
    p.sq1mr2 = sqrtf(1.0f - p.rho * p.rho);
  }

  // jnp.maximum: NaN propagates
  __device__ static float jmax(float a, float b) { return (a > b || a != a) ? a : b; }

  __device__ static void rinit_x(const ModelDesc&, const P& p, Key, float, float* x, const float*) {
    x[0] = p.V0;
    x[1] = 1105.0f;
  }

  __device__ static void draw_x(const P&, Key k, uint64_t i, float* z) { z[0] = normal_at(k, i); }

  __device__ static void step_z_x(const ModelDesc&, const P& p, const float* z, float, float, float* x,
                                  const float* cv) {
    const float V = x[0], S = x[1], y_prev = cv[0];
    const float sV = sqrtf(V);
    const float dWs = ((y_prev - p.mu) + 0.5f * V) / sV;
    const float dWv = p.rho * dWs + p.sq1mr2 * z[0];
    const float Sn = S + S * (p.mu + sqrtf(jmax(V, 1e-32f)) * dWs);
    const float Vn = (V + p.kappa * (p.theta - V)) + (p.xi * sV) * dWv;
    x[0] = jmax(Vn, 1e-32f);
    x[1] = Sn;
  }

  __device__ static void step_x(const ModelDesc& md, const P& p, Key k, uint64_t i, float t, float dt, float* x,
                                const float* cv) {
    float z[1];
    draw_x(p, k, i, z);
    step_z_x(md, p, z, t, dt, x, cv);
  }

  __device__ static float dmeas_x(const ModelDesc&, const P& p, const float* y, const float* x, float,
                                  const float*) {
    const float V = x[0];
    const float loc = p.mu - 0.5f * V;
    const float s = sqrtf(V);
    const float s2 = s * s;
    const float r = y[0] - loc;
    return (d_log(6.28318548f * s2) + (r * r) / s2) / -2.0f;
  }

  __device__ static void rmeas_x(const ModelDesc&, const P& p, Key k, const float* x, float, float* y,
                                 const float*) {
    const float V = x[0];
    y[0] = (p.mu - 0.5f * V) + sqrtf(V) * normal_at(k, 0);
  }
};

}  // namespace mp
