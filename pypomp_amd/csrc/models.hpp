// Device model functors: the native form of the reference's rinit / rproc /
// dmeas plugins (pypomp/core/model_mechanics.py:361-697).  Each model draws its
// random numbers from the particle key it is handed, in the order documented
// on the model, so that the stream matches the reference's per-particle
// vmapped plugin (model_mechanics.py:584-588).
#pragma once

#include "common.hpp"

namespace mp {

enum ModelKind : int { MODEL_LG = 1, MODEL_SEIR = 2, MODEL_SV = 3, MODEL_SPX = 4 };

// Host-filled description of a model instance (plain data, passed by value).
struct ModelDesc {
  int kind;
  int S;            // state words per particle
  int d;            // parameters
  int dy;           // observation words
  int dx;           // LG: state dimension
  int lg_sep_init;  // LG: separate initial Cholesky factor P follows X0
  uint32_t accum_mask;
  int identity_trans;  // estimation scale == natural scale
  int vectorized;      // @vectorized rproc: one draw per sub-step key, particle j at index j
};

__device__ __forceinline__ int tri(int d) { return d * (d + 1) / 2; }

// ---------------------------------------------------------------------------
// Linear Gaussian (pypomp/models/linear_gaussian.py:139-191), dx, dy <= MD.
// theta = A (dx*dx, row-major) | C (dy*dx) | Q lower-tri Cholesky entries |
//         R lower-tri Cholesky entries | X0 (dx) | [P lower-tri]  (natural scale).
// rinit: x = X0 + L0 z, z = normal(key, (dx,))  (L0 = L_Q unless P is given)
// step:  x = A x + L_Q z, z = normal(key, (dx,))
// dmeas: log N(y; C x, L_R L_R^T), evaluated as the reference does through the
//        Cholesky factor of the covariance (jax.scipy.stats.multivariate_normal).
// Estimation scale: log of the diagonal Q and R entries.
// All loops run to the compile-time bound MD with a guard; with EXACT the
// dimensions are the constant MD, so every array index folds to a constant.
// ---------------------------------------------------------------------------
#define MP_UNROLL _Pragma("unroll")

template <int MD, bool EXACT>
struct LGT {
  static constexpr int MAXS = MD;
  static constexpr bool FIXED_S = EXACT;  // S == MAXS for every instance
  static constexpr int MAXP = 2 * MD * MD + 3 * (MD * (MD + 1) / 2) + MD;
  // the MOP adjoint covers dx, dy <= 2 (the gradient accumulators of larger
  // shapes would not fit in registers)
  static constexpr bool DIFFERENTIABLE = MD <= 2;
  struct P {
    float A[MD][MD], C[MD][MD], LQ[MD][MD], L0[MD][MD], LC[MD][MD], LR[MD][MD];
    float X0[MD];
    float logdet, half_n_log2pi;
    int dx, dy;
  };
  // The dimensions: the constant MD when the size is exact, so that every
  // guard and index below folds at compile time.
  __host__ __device__ static int DX(const P& p) { return EXACT ? MD : p.dx; }
  __host__ __device__ static int DY(const P& p) { return EXACT ? MD : p.dy; }

  __host__ __device__ static void from_est(const ModelDesc& md, float* th) {
    const int dx = EXACT ? MD : md.dx, dy = EXACT ? MD : md.dy;
    // diagonal entries of the Q and R triangles
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < dx) {
      const int o = dx * dx + dy * dx + i * (i + 1) / 2 + i;
      th[o] = d_exp(th[o]);
    }
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < dy) {
      const int o = dx * dx + dy * dx + dx * (dx + 1) / 2 + i * (i + 1) / 2 + i;
      th[o] = d_exp(th[o]);
    }
  }

  __device__ static void prepare(const ModelDesc& md, const float* th, P& p) {
    const int dx = EXACT ? MD : md.dx, dy = EXACT ? MD : md.dy;
    p.dx = dx;
    p.dy = dy;
    MP_UNROLL for (int i = 0; i < MD; ++i) MP_UNROLL for (int j = 0; j < MD; ++j) {
      p.A[i][j] = p.C[i][j] = p.LQ[i][j] = p.L0[i][j] = p.LC[i][j] = 0.f;
    }
    const int oC = dx * dx, oQ = oC + dy * dx, oR = oQ + dx * (dx + 1) / 2, oX = oR + dy * (dy + 1) / 2,
              oP = oX + dx;
    float LR[MD][MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) MP_UNROLL for (int j = 0; j < MD; ++j) {
      if (i < dx && j < dx) p.A[i][j] = th[i * dx + j];
      if (i < dy && j < dx) p.C[i][j] = th[oC + i * dx + j];
      p.LQ[i][j] = (i < dx && j <= i) ? th[oQ + i * (i + 1) / 2 + j] : 0.f;
      LR[i][j] = (i < dy && j <= i) ? th[oR + i * (i + 1) / 2 + j] : 0.f;
      p.LR[i][j] = LR[i][j];
      if (md.lg_sep_init) p.L0[i][j] = (i < dx && j <= i) ? th[oP + i * (i + 1) / 2 + j] : 0.f;
      else p.L0[i][j] = p.LQ[i][j];
    }
    MP_UNROLL for (int i = 0; i < MD; ++i) p.X0[i] = i < dx ? th[oX + i] : 0.f;
    // Cholesky factor of cov = LR LR^T (float32) and its log determinant.
    float cov[MD][MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) MP_UNROLL for (int j = 0; j < MD; ++j) {
      float s = 0.f;
      MP_UNROLL for (int k = 0; k < MD; ++k) s += LR[i][k] * LR[j][k];
      cov[i][j] = s;
    }
    float logdet = 0.f;
    MP_UNROLL for (int j = 0; j < MD; ++j) if (j < dy) {
      float s = cov[j][j];
      MP_UNROLL for (int k = 0; k < MD; ++k) if (k < j) s -= p.LC[j][k] * p.LC[j][k];
      const float djj = sqrtf(s);
      p.LC[j][j] = djj;
      logdet += d_log(djj);
      MP_UNROLL for (int i = 0; i < MD; ++i) if (i > j && i < dy) {
        float t = cov[i][j];
        MP_UNROLL for (int k = 0; k < MD; ++k) if (k < j) t -= p.LC[i][k] * p.LC[j][k];
        p.LC[i][j] = t / djj;
      }
    }
    p.logdet = logdet;
    p.half_n_log2pi = 0.5f * (float)dy * 1.83787707f;
  }

  __device__ static void rinit(const ModelDesc&, const P& p, Key k, float, float* x) {
    float z[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) z[i] = i < DX(p) ? normal_at(k, i) : 0.f;
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DX(p)) {
      float s = p.L0[i][0] * z[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) if (j < DX(p)) s += p.L0[i][j] * z[j];
      x[i] = p.X0[i] + s;
    }
  }

  // The step's noise depends on the particle's key only, not on its state:
  // kernels may draw it before the ancestor is known (draw + step_z == step).
  static constexpr int NOISE = MD;
  // the draw reads the parameters (their dx) unless the size is exact
  static constexpr bool DRAW_USES_P = !EXACT;
  static constexpr bool HOIST = false;  // no particle-independent step or density terms
  __device__ static void draw(const P& p, Key k, float* z) {
    MP_UNROLL for (int i = 0; i < MD; ++i) z[i] = i < DX(p) ? normal_at(k, i) : 0.f;
  }

  __device__ static void step(const ModelDesc& md, const P& p, Key k, float t, float dt, float* x) {
    float z[MD];
    draw(p, k, z);
    step_z(md, p, z, t, dt, x);
  }

  __device__ static void step_z(const ModelDesc&, const P& p, const float* z, float, float, float* x) {
    float xn[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      float a = p.A[i][0] * x[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) if (j < DX(p)) a += p.A[i][j] * x[j];
      float s = p.LQ[i][0] * z[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) if (j < DX(p)) s += p.LQ[i][j] * z[j];
      xn[i] = a + s;
    }
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DX(p)) x[i] = xn[i];
  }

  __device__ static float dmeas(const ModelDesc&, const P& p, const float* y, const float* x, float) {
    float r[MD], s[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      float m = p.C[i][0] * x[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) if (j < DX(p)) m += p.C[i][j] * x[j];
      r[i] = i < DY(p) ? y[i] - m : 0.f;
    }
    float quad = 0.f;
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DY(p)) {
      float a = r[i];
      MP_UNROLL for (int k = 0; k < MD; ++k) if (k < i) a -= p.LC[i][k] * s[k];
      s[i] = a / p.LC[i][i];
      quad += s[i] * s[i];
    }
    return -0.5f * quad - p.half_n_log2pi - p.logdet;
  }

  // y = C x + LR z, z = normal(key, (dy,)) (linear_gaussian.py:181-191)
  __device__ static void rmeas(const ModelDesc&, const P& p, Key k, const float* x, float, float* y) {
    float z[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) z[i] = i < DY(p) ? normal_at(k, i) : 0.f;
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DY(p)) {
      float m = p.C[i][0] * x[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) if (j < DX(p)) m += p.C[i][j] * x[j];
      float s = p.LR[i][0] * z[0];
      MP_UNROLL for (int j = 1; j < MD; ++j) if (j < DY(p)) s += p.LR[i][j] * z[j];
      y[i] = m + s;
    }
  }

  // ---- derivatives for the MOP adjoint (natural-scale parameters) ----
  // theta offsets: A, C, LQ (lower triangle), LR (lower triangle), X0, [L0]
  __device__ static int oC(const P& p) { return DX(p) * DX(p); }
  __device__ static int oQ(const P& p) { return oC(p) + DY(p) * DX(p); }
  __device__ static int oR(const P& p) { return oQ(p) + DX(p) * (DX(p) + 1) / 2; }
  __device__ static int oX(const P& p) { return oR(p) + DY(p) * (DY(p) + 1) / 2; }
  __device__ static int oP(const P& p) { return oX(p) + DX(p); }

  __device__ static float dmeas_aux(const P&, const float*) { return 0.f; }
  __device__ static float dmeas_a(const ModelDesc& md, const P& p, const float* y, const float* x, float) {
    return dmeas(md, p, y, x, 0.f);
  }

  // log N(y; C x, S), S = LR LR^T = LC LC^T:  with w = S^-1 (y - C x),
  //   d/dx = C^T w,  d/dC = w x^T,  d/dS = (w w^T - S^-1) / 2,  d/dLR = 2 (d/dS) LR.
  __device__ static void dmeas_grad(const ModelDesc&, const P& p, const float* y, const float* x, float,
                                    float wgt, float* gx, double* gth) {
    float r[MD], s[MD], w[MD], Li[MD][MD], Si[MD][MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      float m = 0.f;
      MP_UNROLL for (int j = 0; j < MD; ++j) if (j < DX(p)) m += p.C[i][j] * x[j];
      r[i] = i < DY(p) ? y[i] - m : 0.f;
    }
    // s = LC^-1 r, w = LC^-T s; Li = LC^-1 (lower), Si = Li^T Li
    MP_UNROLL for (int i = 0; i < MD; ++i) {
      float a = r[i];
      MP_UNROLL for (int k = 0; k < MD; ++k) if (k < i) a -= p.LC[i][k] * s[k];
      s[i] = i < DY(p) ? a / p.LC[i][i] : 0.f;
    }
    MP_UNROLL for (int ii = MD - 1; ii >= 0; --ii) {
      float a = s[ii];
      MP_UNROLL for (int k = 0; k < MD; ++k) if (k > ii && k < DY(p)) a -= p.LC[k][ii] * w[k];
      w[ii] = ii < DY(p) ? a / p.LC[ii][ii] : 0.f;
    }
    MP_UNROLL for (int c = 0; c < MD; ++c) MP_UNROLL for (int i = 0; i < MD; ++i) {
      float a = (i == c) ? 1.f : 0.f;
      MP_UNROLL for (int k = 0; k < MD; ++k) if (k < i) a -= p.LC[i][k] * Li[k][c];
      Li[i][c] = (i < DY(p) && c < DY(p)) ? a / p.LC[i][i] : 0.f;
    }
    MP_UNROLL for (int i = 0; i < MD; ++i) MP_UNROLL for (int j = 0; j < MD; ++j) {
      float a = 0.f;
      MP_UNROLL for (int k = 0; k < MD; ++k) a += Li[k][i] * Li[k][j];
      Si[i][j] = a;
    }
    MP_UNROLL for (int j = 0; j < MD; ++j) {
      float a = 0.f;
      MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DY(p)) a += p.C[i][j] * w[i];
      gx[j] = a;
    }
    const double wd = (double)wgt;
    const int oc = oC(p), orr = oR(p);
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DY(p))
      MP_UNROLL for (int j = 0; j < MD; ++j) if (j < DX(p)) gth[oc + i * DX(p) + j] += wd * (double)(w[i] * x[j]);
    // d/dLR_ij = sum_k (w_i w_k - Si_ik) LR_kj, j <= i
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DY(p)) MP_UNROLL for (int j = 0; j <= i; ++j) {
      float a = 0.f;
      MP_UNROLL for (int k = 0; k < MD; ++k) if (k < DY(p)) a += (w[i] * w[k] - Si[i][k]) * p.LR[k][j];
      gth[orr + i * (i + 1) / 2 + j] += wd * (double)a;
    }
  }

  // x_out = A x_in + LQ z, z redrawn from the particle key
  __device__ static void step_vjp(const ModelDesc&, const P& p, Key k, const float* xin, const float*,
                                  const float* xbar_out, float* xbar_in, double* gth) {
    float z[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) z[i] = i < DX(p) ? normal_at(k, i) : 0.f;
    MP_UNROLL for (int j = 0; j < MD; ++j) {
      float a = 0.f;
      MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DX(p)) a += p.A[i][j] * xbar_out[i];
      xbar_in[j] = a;
    }
    const int oq = oQ(p);
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DX(p)) {
      const double b = (double)xbar_out[i];
      MP_UNROLL for (int j = 0; j < MD; ++j) if (j < DX(p)) gth[i * DX(p) + j] += b * (double)xin[j];
      MP_UNROLL for (int j = 0; j <= i; ++j) gth[oq + i * (i + 1) / 2 + j] += b * (double)z[j];
    }
  }

  // x_0 = X0 + L0 z (L0 = LQ unless a separate initial factor is given)
  __device__ static void rinit_vjp(const ModelDesc& md, const P& p, Key k, float, const float* xbar, double* gth) {
    float z[MD];
    MP_UNROLL for (int i = 0; i < MD; ++i) z[i] = i < DX(p) ? normal_at(k, i) : 0.f;
    const int ox = oX(p), ol = md.lg_sep_init ? oP(p) : oQ(p);
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < DX(p)) {
      const double b = (double)xbar[i];
      gth[ox + i] += b;
      MP_UNROLL for (int j = 0; j <= i; ++j) gth[ol + i * (i + 1) / 2 + j] += b * (double)z[j];
    }
  }

  // d(natural)/d(estimation): exp on the diagonals of the LQ and LR triangles
  __device__ static void from_est_jac(const ModelDesc& md, const float* est, float* jac) {
    const int dx = EXACT ? MD : md.dx, dy = EXACT ? MD : md.dy;
    MP_UNROLL for (int q = 0; q < MAXP; ++q) jac[q] = 1.f;
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < dx) {
      const int o = dx * dx + dy * dx + i * (i + 1) / 2 + i;
      jac[o] = d_exp(est[o]);
    }
    MP_UNROLL for (int i = 0; i < MD; ++i) if (i < dy) {
      const int o = dx * dx + dy * dx + dx * (dx + 1) / 2 + i * (i + 1) / 2 + i;
      jac[o] = d_exp(est[o]);
    }
  }
};

// dx = dy = 1 and dx = dy = 2 get fully constant indexing; other shapes use
// a generic instance with run-time dimensions: up to 4, or up to 8 (the
// prepared parameters then take 1.6 KB, in LDS where the particles share them).
using LG1 = LGT<1, true>;
using LG2 = LGT<2, true>;
using LG4 = LGT<4, false>;
using LG8 = LGT<8, false>;
constexpr int kLgMaxDim = 8;

// ---------------------------------------------------------------------------
// Stochastic volatility with leverage (the MOP workload).  State (v, eta):
//   v_0  = V0 + S0 z,                 eta_0 = 0,         z = normal(key, (1,))
//   v_t  = mu + phi (v_{t-1} - mu) + sigma eta_t,   eta_t = normal(key, (1,))
//   r_t | v_t, eta_t ~ N(exp(v_t/2) rho eta_t, exp(v_t) (1 - rho^2))
// i.e. corr(eta_t, eps_t) = rho for r_t = exp(v_t/2) eps_t.
// theta = (mu, phi, sigma, rho, V0, S0); estimation scale: phi = tanh, sigma = exp,
// rho = tanh, S0 = exp, mu and V0 unchanged.
// ---------------------------------------------------------------------------
struct SV {
  static constexpr int MAXS = 2;
  static constexpr bool FIXED_S = true;
  static constexpr int MAXP = 6;
  struct P {
    float mu, phi, sigma, rho, V0, S0;
    float c0;       // -0.5 log(2 pi) - 0.5 log(1 - rho^2)
    float inv1mr2;  // 1 / (1 - rho^2)
  };

  __host__ __device__ static void from_est(const ModelDesc&, float* th) {
    th[1] = d_tanh(th[1]);
    th[2] = d_exp(th[2]);
    th[3] = d_tanh(th[3]);
    th[5] = d_exp(th[5]);
  }

  __device__ static void prepare(const ModelDesc&, const float* th, P& p) {
    p.mu = th[0];
    p.phi = th[1];
    p.sigma = th[2];
    p.rho = th[3];
    p.V0 = th[4];
    p.S0 = th[5];
    float omr2 = 1.0f - p.rho * p.rho;
    p.inv1mr2 = 1.0f / omr2;
    p.c0 = -0.918938533f - 0.5f * d_log(omr2);
  }

  __device__ static void rinit(const ModelDesc&, const P& p, Key k, float, float* x) {
    float z = normal_at(k, 0);
    x[0] = p.V0 + p.S0 * z;
    x[1] = 0.0f;
  }

  // the step's one normal depends on the particle's key only (draw + step_z == step)
  static constexpr int NOISE = 1;
  static constexpr bool DRAW_USES_P = false;
  static constexpr bool HOIST = false;
  __device__ static void draw(const P&, Key k, float* z) { z[0] = normal_at(k, 0); }

  __device__ static void step_z(const ModelDesc&, const P& p, const float* zp, float, float, float* x) {
    const float z = zp[0];
    x[0] = p.mu + p.phi * (x[0] - p.mu) + p.sigma * z;
    x[1] = z;
  }

  __device__ static void step(const ModelDesc& md, const P& p, Key k, float t, float dt, float* x) {
    float z[1];
    draw(p, k, z);
    step_z(md, p, z, t, dt, x);
  }

  // exp(-v/2): the one transcendental of the density, kept by the adjoint's
  // recompute pass for the backward sweep
  __device__ static float dmeas_aux(const P&, const float* x) { return d_exp(-0.5f * x[0]); }

  __device__ static float dmeas_a(const ModelDesc&, const P& p, const float* y, const float* x, float e) {
    float u = y[0] * e - p.rho * x[1];
    return p.c0 - 0.5f * x[0] - 0.5f * u * u * p.inv1mr2;
  }

  __device__ static float dmeas(const ModelDesc& md, const P& p, const float* y, const float* x, float) {
    return dmeas_a(md, p, y, x, dmeas_aux(p, x));
  }

  // r = exp(v/2) (rho eta + sqrt(1 - rho^2) e), e = normal(key, (1,))
  __device__ static void rmeas(const ModelDesc&, const P& p, Key k, const float* x, float, float* y) {
    const float e = normal_at(k, 0);
    const float sd = sqrtf(1.0f - p.rho * p.rho);
    y[0] = d_exp(0.5f * x[0]) * (p.rho * x[1] + sd * e);
  }

  // ---- derivatives for the MOP adjoint (natural-scale parameters) ----
  static constexpr bool DIFFERENTIABLE = true;

  // d(dmeas)/dx into gx[S], accumulated d(dmeas)/dtheta into gth[d] scaled by
  // w; e = dmeas_aux(x).
  __device__ static void dmeas_grad(const ModelDesc&, const P& p, const float* y, const float* x, float e, float w,
                                    float* gx, double* gth) {
    const float ry = y[0] * e;
    const float u = ry - p.rho * x[1];
    const float k = p.inv1mr2;
    gx[0] = -0.5f + 0.5f * u * ry * k;
    gx[1] = u * p.rho * k;
    gth[3] += (double)w * (double)(p.rho * k + u * x[1] * k - u * u * p.rho * k * k);
  }

  // Vector-Jacobian product of one sub-step x_out = step(x_in): given xbar_out,
  // writes xbar_in and accumulates theta-bar.  The step's normal is x_out[1]
  // (eta), so nothing is redrawn.
  __device__ static void step_vjp(const ModelDesc&, const P& p, Key, const float* xin, const float* xout,
                                  const float* xbar_out, float* xbar_in, double* gth) {
    const float z = xout[1];
    const double vb = xbar_out[0];
    xbar_in[0] = p.phi * xbar_out[0];
    xbar_in[1] = 0.f;
    gth[0] += vb * (double)(1.0f - p.phi);
    gth[1] += vb * (double)(xin[0] - p.mu);
    gth[2] += vb * (double)z;
  }

  __device__ static void rinit_vjp(const ModelDesc&, const P&, Key k, float, const float* xbar, double* gth) {
    const float z = normal_at(k, 0);
    gth[4] += (double)xbar[0];
    gth[5] += (double)xbar[0] * (double)z;
  }

  // d(natural)/d(estimation) for each parameter (diagonal transform), at est.
  __device__ static void from_est_jac(const ModelDesc&, const float* est, float* jac) {
    const float ph = d_tanh(est[1]), rh = d_tanh(est[3]);
    jac[0] = 1.f;
    jac[1] = 1.f - ph * ph;
    jac[2] = d_exp(est[2]);
    jac[3] = 1.f - rh * rh;
    jac[4] = 1.f;
    jac[5] = d_exp(est[5]);
  }
};

}  // namespace mp
