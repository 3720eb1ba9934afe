/*
 * cpu_pf.c -- C restatement of the particle-filter hot path for the CPU
 * baseline.  TEST INFRASTRUCTURE ONLY (built by oracle/Makefile into
 * oracle/_ref/libcpupf.so, loaded by tests and by bench.py's cpu_baseline leg).
 *
 * Restates, like the NumPy oracle (oracle/filters.py):
 *  - the JAX threefry key tree (split / random_bits / uniform / normal),
 *  - the bootstrap filter _pfilter_internal / _pfilter_step
 *    (pypomp/core/algorithms/pfilter.py:15-182) with thresh = 0,
 *  - the MOP(alpha) value _mop_internal (core/algorithms/mop.py:19-135),
 *  - the systematic resampler with fixed-point prefix sums (oracle/resample.py),
 * for the linear-Gaussian model (dx = dy <= 4, separate initial factor) and the
 * stochastic-volatility model.  float32 state arithmetic, float64 log-likelihood
 * sum.  OpenMP runs one replicate filter per thread, as the reference's CPU
 * sharding does (core/algorithms/helpers.py:494-527, 650-708).
 * Must be compiled with -ffp-contract=off (the resampler's exp uses separate
 * multiplies and adds).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint32_t k0, k1;
} key_t2;

static inline uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static void threefry(uint32_t k0, uint32_t k1, uint32_t* x0p, uint32_t* x1p) {
  static const int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  uint32_t ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  uint32_t x0 = *x0p + ks[0], x1 = *x1p + ks[1];
  for (int blk = 0; blk < 5; ++blk) {
    for (int r = 0; r < 4; ++r) {
      x0 += x1;
      x1 = rotl(x1, R[(blk & 1) * 4 + r]);
      x1 ^= x0;
    }
    x0 += ks[(blk + 1) % 3];
    x1 += ks[(blk + 2) % 3] + (uint32_t)(blk + 1);
  }
  *x0p = x0;
  *x1p = x1;
}

static inline key_t2 split_at(key_t2 k, uint64_t i) {
  uint32_t a = (uint32_t)(i >> 32), b = (uint32_t)i;
  threefry(k.k0, k.k1, &a, &b);
  key_t2 o = {a, b};
  return o;
}

static inline uint32_t bits_at(key_t2 k, uint64_t i) {
  uint32_t a = (uint32_t)(i >> 32), b = (uint32_t)i;
  threefry(k.k0, k.k1, &a, &b);
  return a ^ b;
}

static inline float unit_float(uint32_t bits) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = (bits >> 9) | 0x3F800000u;
  return c.f - 1.0f;
}

static inline float uniform01(uint32_t bits) {
  float f = unit_float(bits);
  return f > 0.f ? f : 0.f;
}


/* ---- deterministic elementary functions (same operation sequences as the
 * engine and oracle/detmath.py; compiled with -ffp-contract=off) ---- */
static inline float bf(uint32_t b) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = b;
  return c.f;
}
static inline uint32_t fb(float f) {
  union {
    uint32_t u;
    float f;
  } c;
  c.f = f;
  return c.u;
}
static inline float pow2f(int k) { return bf((uint32_t)(k + 127) << 23); }

static float d_log(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  int e = 0;
  if (x < 1.17549435e-38f) {
    x = x * 8388608.0f;
    e = -23;
  }
  const uint32_t b = fb(x);
  e += (int)((b >> 23) & 0xffu) - 127;
  float m = bf((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f, s = f / (2.0f + f), z = s * s;
  float p = 0.111111112f;
  p = p * z + 0.142857149f;
  p = p * z + 0.2f;
  p = p * z + 0.333333343f;
  p = p * z + 1.0f;
  const float lm = (2.0f * s) * p, fe = (float)e;
  return fe * 0.693145751953125f + (lm + fe * 1.42860677e-06f);
}

static float d_log1p(float x) {
  if (x == 0.0f) return x;
  const float u = 1.0f + x;
  if (u == 1.0f) return x;
  return d_log(u) * (x / (u - 1.0f));
}

static float d_exp(float x) {
  if (x != x) return x;
  if (x > 88.7228394f) return INFINITY;
  if (x < -87.3365479f) return 0.0f;
  const float k = rintf(x * 1.44269502f);
  const float r = (x - k * 0.693145751953125f) - k * 1.42860677e-06f;
  float p = 1.98412701e-04f;
  p = p * r + 1.38888892e-03f;
  p = p * r + 8.33333377e-03f;
  p = p * r + 4.16666679e-02f;
  p = p * r + 0.166666672f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  const int ki = (int)k;
  if (ki < -126) return (p * pow2f(ki + 64)) * 5.42101086e-20f;
  if (ki > 127) return (p * pow2f(ki - 64)) * 1.84467441e+19f;
  return p * pow2f(ki);
}

static float d_expm1(float x) {
  if (x == 0.0f) return x;
  const float u = d_exp(x);
  if (u == 1.0f) return x;
  const float um1 = u - 1.0f;
  if (um1 == -1.0f) return -1.0f;
  if (u == INFINITY) return u;
  return um1 * (x / d_log(u));
}

static float d_tanh(float x) {
  const float ax = x < 0.0f ? -x : x;
  const float t = d_expm1(-2.0f * ax);
  const float r = -t / (t + 2.0f);
  return x < 0.0f ? -r : r;
}

static inline float erfinv_f32(float x) {
  float w = -d_log((1.0f - x) * (1.0f + x)), p;
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
  float r = p * x;
  if (fabsf(x) == 1.0f) r = x * 3.40282347e+38f;
  return r;
}

static inline float normal_at(key_t2 k, uint64_t i) {
  const float lo = -0.99999994f;
  float u = unit_float(bits_at(k, i)) * 2.0f + lo;
  if (!(u > lo)) u = lo;
  return 1.41421354f * erfinv_f32(u);
}

/* ---- fixed-point systematic resampler (oracle/resample.py) ---- */
static const double EXPC[14] = {0x1p+0, 0x1p+0, 0x1p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
                                0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
                                0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22,
                                0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33};

static int shift_for(int64_t J) {
  int b = 0;
  uint64_t v = (uint64_t)J;
  while (v) {
    ++b;
    v >>= 1;
  }
  return 62 - b;
}

static uint64_t fixed_weight(double d, int sh) {
  const double cutoff = -(double)(sh + 2) * 0.6931471805599453;
  if (!(d >= cutoff)) return 0;
  double k = rint(d * 1.4426950408889634);
  double r = d - k * 6.93147180369123816490e-01;
  r = r - k * 1.90821492927058770002e-10;
  double p = EXPC[13];
  for (int n = 12; n >= 0; --n) p = p * r + EXPC[n];
  return (uint64_t)ldexp(p, (int)k + sh);
}

static uint64_t comb_threshold(float u, int64_t j, int64_t J, uint64_t total) {
  float f = (u + (float)j) / (float)J;
  union {
    float f;
    uint32_t u;
  } c;
  c.f = f;
  uint32_t e = (c.u >> 23) & 0xffu;
  if (e == 0) return 0;
  uint64_t mant = (uint64_t)((c.u & 0x7fffffu) | 0x800000u);
  int sh = 150 - (int)e;
  uint64_t p_lo = mant * (total & 0xffffffffull), p_hi = mant * (total >> 32);
  uint64_t X = p_hi + (p_lo >> 32), lo32 = p_lo & 0xffffffffull;
  if (sh >= 32) {
    int s = sh - 32;
    return s >= 64 ? 0 : (X >> s);
  }
  return (X << (32 - sh)) + (lo32 >> sh);
}

/* log of a positive double: the deterministic sequence of oracle/detmath.py dd_log */
static double dd_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  union {
    double d;
    uint64_t u;
  } c;
  int e = 0;
  if (x < 2.2250738585072014e-308) {
    x = x * 4503599627370496.0;
    e = -52;
  }
  c.d = x;
  e += (int)((c.u >> 52) & 0x7ffu) - 1023;
  c.u = (c.u & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m = c.d;
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0, s = f / (2.0 + f), z = s * s;
  static const double den[10] = {21.0, 19.0, 17.0, 15.0, 13.0, 11.0, 9.0, 7.0, 5.0, 3.0};
  double p = 1.0 / 23.0;
  for (int i = 0; i < 10; ++i) p = p * z + 1.0 / den[i];
  p = p * z + 1.0;
  const double lm = (2.0 * s) * p, de = (double)e;
  return de * 6.93147180369123816490e-01 + (lm + de * 1.90821492927058770002e-10);
}

/* inclusive prefix sums C of the fixed-point weights of lw (max m); returns the total */
static uint64_t fixed_prefix(const float* lw, float m, int64_t J, uint64_t* C) {
  const int sh = shift_for(J);
  uint64_t run = 0;
  for (int64_t i = 0; i < J; ++i) {
    uint64_t e;
    if (m == -INFINITY) e = 1;
    else if (m == INFINITY) e = lw[i] == INFINITY;
    else e = fixed_weight((double)lw[i] - (double)m, sh);
    run += e;
    C[i] = run;
  }
  return run;
}

/* log-likelihood term m + log(total 2^-sh) (common.hpp step_loglik) */
static double step_loglik(float m, uint64_t total, int64_t J) {
  if (m == -INFINITY) return -INFINITY;
  if (!(m < INFINITY)) return NAN;
  return (double)m + dd_log(ldexp((double)total, -shift_for(J)));
}

/* ancestors of J particles from prefix sums C (total) and uniform u */
static void resample(const uint64_t* C, uint64_t total, int64_t J, float u, int32_t* anc) {
  int64_t i = 0;
  for (int64_t j = 0; j < J; ++j) { /* merge: thresholds are non-decreasing in j */
    uint64_t t = comb_threshold(u, j, J, total);
    while (i < J && C[i] <= t) ++i;
    anc[j] = (int32_t)(i < J ? i : J - 1);
  }
}

/* ---- models ---- */
typedef struct {
  int dx;
  float A[4][4], LQ[4][4], L0[4][4], C[4][4], LC[4][4], X0[4], logdet;
} lg_t;

static void lg_prepare(const float* th, int dx, lg_t* p) {
  memset(p, 0, sizeof *p);
  p->dx = dx;
  int o = 0;
  for (int i = 0; i < dx; ++i)
    for (int j = 0; j < dx; ++j) p->A[i][j] = th[o++];
  for (int i = 0; i < dx; ++i)
    for (int j = 0; j < dx; ++j) p->C[i][j] = th[o++];
  for (int i = 0; i < dx; ++i)
    for (int j = 0; j <= i; ++j) p->LQ[i][j] = th[o++];
  float LR[4][4] = {{0}};
  for (int i = 0; i < dx; ++i)
    for (int j = 0; j <= i; ++j) LR[i][j] = th[o++];
  for (int i = 0; i < dx; ++i) p->X0[i] = th[o++];
  for (int i = 0; i < dx; ++i)
    for (int j = 0; j <= i; ++j) p->L0[i][j] = th[o++];
  float cov[4][4];
  for (int i = 0; i < dx; ++i)
    for (int j = 0; j < dx; ++j) {
      float s = 0.f;
      for (int k = 0; k < dx; ++k) s += LR[i][k] * LR[j][k];
      cov[i][j] = s;
    }
  float ld = 0.f;
  for (int j = 0; j < dx; ++j) {
    float s = cov[j][j];
    for (int k = 0; k < j; ++k) s -= p->LC[j][k] * p->LC[j][k];
    float d = sqrtf(s);
    p->LC[j][j] = d;
    ld += d_log(d);
    for (int i = j + 1; i < dx; ++i) {
      float t = cov[i][j];
      for (int k = 0; k < j; ++k) t -= p->LC[i][k] * p->LC[j][k];
      p->LC[i][j] = t / d;
    }
  }
  p->logdet = ld;
}

/* Bootstrap filter of the LG model with separate initial factor (thresh = 0):
 * returns the FP64 log-likelihood; X (dx*J) and scratch provided by the caller. */
static double lg_pfilter_one(const float* th, int dx, const float* ys, int T, int64_t J, key_t2 key, float* X,
                             float* Xn, float* lw, uint64_t* C, int32_t* anc) {
  lg_t p;
  lg_prepare(th, dx, &p);
  key_t2 run = split_at(key, 0);
  for (int64_t j = 0; j < J; ++j) {
    key_t2 pk = split_at(key, 1 + (uint64_t)j);
    float z[4];
    for (int i = 0; i < dx; ++i) z[i] = normal_at(pk, i);
    for (int i = 0; i < dx; ++i) {
      float s = p.L0[i][0] * z[0];
      for (int k = 1; k < dx; ++k) s += p.L0[i][k] * z[k];
      X[i * J + j] = p.X0[i] + s;
    }
  }
  const float logJ = d_log((float)J);
  float wconst = d_log(1.0f / (float)J);
  double ll = 0.0;
  const float hl2p = (0.5f * (float)dx) * 1.83787707f;
  for (int t = 0; t < T; ++t) {
    key_t2 sub = split_at(run, 1);
    run = split_at(run, 0);
    key_t2 sk = split_at(sub, 1);
    const float* y = ys + (int64_t)t * dx;
    float m = -INFINITY;
    for (int64_t j = 0; j < J; ++j) {
      int64_t s0 = t == 0 ? j : anc[j];
      float x[4], z[4], xn[4];
      for (int i = 0; i < dx; ++i) x[i] = X[i * J + s0];
      key_t2 pk = split_at(sk, (uint64_t)j);
      for (int i = 0; i < dx; ++i) z[i] = normal_at(pk, i);
      for (int i = 0; i < dx; ++i) {
        float a = p.A[i][0] * x[0], s = p.LQ[i][0] * z[0];
        for (int k = 1; k < dx; ++k) {
          a += p.A[i][k] * x[k];
          s += p.LQ[i][k] * z[k];
        }
        xn[i] = a + s;
      }
      float r[4], sv[4], quad = 0.f;
      for (int i = 0; i < dx; ++i) {
        float mm = p.C[i][0] * xn[0];
        for (int k = 1; k < dx; ++k) mm += p.C[i][k] * xn[k];
        r[i] = y[i] - mm;
      }
      for (int i = 0; i < dx; ++i) {
        float a2 = r[i];
        for (int k = 0; k < i; ++k) a2 -= p.LC[i][k] * sv[k];
        sv[i] = a2 / p.LC[i][i];
        quad += sv[i] * sv[i];
      }
      const float w = wconst + ((-0.5f * quad - hl2p) - p.logdet);
      for (int i = 0; i < dx; ++i) Xn[i * J + j] = xn[i];
      lw[j] = w;
      if (w > m) m = w;
    }
    const uint64_t total = fixed_prefix(lw, m, J, C);
    ll += step_loglik(m, total, J);
    key_t2 rk = split_at(run, 1);
    run = split_at(run, 0);
    if (t < T - 1) resample(C, total, J, uniform01(bits_at(rk, 0)), anc);
    wconst = 0.0f - logJ;
    float* tmp = X;
    X = Xn;
    Xn = tmp;
  }
  return ll;
}

/* R independent LG filters (theta shared), one per OpenMP thread. */
int cpu_lg_pfilter(const float* theta, int dx, const float* ys, int T, int64_t J, const uint32_t* keys, int R,
                   double* loglik) {
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < R; ++r) {
    float* X = (float*)malloc(sizeof(float) * dx * J);
    float* Xn = (float*)malloc(sizeof(float) * dx * J);
    float* lw = (float*)malloc(sizeof(float) * J);
    uint64_t* C = (uint64_t*)malloc(sizeof(uint64_t) * J);
    int32_t* anc = (int32_t*)malloc(sizeof(int32_t) * J);
    if (!X || !Xn || !lw || !C || !anc) {
      err = 1;
    } else {
      key_t2 k = {keys[2 * r], keys[2 * r + 1]};
      loglik[r] = lg_pfilter_one(theta, dx, ys, T, J, k, X, Xn, lw, C, anc);
    }
    free(X);
    free(Xn);
    free(lw);
    free(C);
    free(anc);
  }
  return err;
}

/* MOP(alpha) value of the SV model (theta on the estimation scale). */
static double sv_mop_one(const float* th_est, const float* ys, int T, int64_t J, float alpha, key_t2 key, float* V,
                         float* Vn, float* E, float* lw, uint64_t* C, int32_t* anc) {
  const float mu = th_est[0], phi = d_tanh(th_est[1]), sig = d_exp(th_est[2]), rho = d_tanh(th_est[3]);
  const float V0 = th_est[4], S0 = d_exp(th_est[5]);
  const float omr2 = 1.0f - rho * rho, k = 1.0f / omr2, c0 = -0.918938533f - 0.5f * d_log(omr2);
  key_t2 run = split_at(key, 0);
  for (int64_t j = 0; j < J; ++j) V[j] = V0 + S0 * normal_at(split_at(key, 1 + (uint64_t)j), 0);
  const double logJ = (double)d_log((float)J);
  double ll = 0.0;
  (void)alpha;
  for (int t = 0; t < T; ++t) {
    key_t2 sub = split_at(run, 1);
    run = split_at(run, 0);
    key_t2 sk = split_at(sub, 1);
    float m = -INFINITY;
    for (int64_t j = 0; j < J; ++j) {
      float v = V[t == 0 ? j : anc[j]];
      float z = normal_at(split_at(sk, (uint64_t)j), 0);
      v = mu + phi * (v - mu) + sig * z;
      float u = ys[t] * d_exp(-0.5f * v) - rho * z;
      float g = c0 - 0.5f * v - 0.5f * u * u * k;
      Vn[j] = v;
      E[j] = z;
      lw[j] = g;
      if (g > m) m = g;
    }
    const uint64_t total = fixed_prefix(lw, m, J, C);
    ll += step_loglik(m, total, J) - logJ;
    key_t2 rk = split_at(run, 1);
    run = split_at(run, 0);
    if (t < T - 1) resample(C, total, J, uniform01(bits_at(rk, 0)), anc);
    float* tmp = V;
    V = Vn;
    Vn = tmp;
  }
  return ll;
}

int cpu_sv_mop(const float* theta_est, const float* ys, int T, int64_t J, float alpha, const uint32_t* keys, int R,
               double* loglik) {
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < R; ++r) {
    float* V = (float*)malloc(sizeof(float) * J);
    float* Vn = (float*)malloc(sizeof(float) * J);
    float* E = (float*)malloc(sizeof(float) * J);
    float* lw = (float*)malloc(sizeof(float) * J);
    uint64_t* C = (uint64_t*)malloc(sizeof(uint64_t) * J);
    int32_t* anc = (int32_t*)malloc(sizeof(int32_t) * J);
    if (!V || !Vn || !E || !lw || !C || !anc) {
      err = 1;
    } else {
      key_t2 k = {keys[2 * r], keys[2 * r + 1]};
      loglik[r] = sv_mop_one(theta_est, ys, T, J, alpha, k, V, Vn, E, lw, C, anc);
    }
    free(V);
    free(Vn);
    free(E);
    free(lw);
    free(C);
    free(anc);
  }
  return err;
}

/* MOP(alpha) value and gradient of the SV model with respect to theta_est, by
 * a hand-written reverse sweep (the derivative jax.grad takes through
 * _mop_internal_mean, core/algorithms/ad_helpers.py:6-8, times -T).  The
 * forward is sv_mop_one's and keeps every step's states v_t, noise z_t,
 * densities m_t and ancestry a_t.  Value-wise the weights wP are one value for
 * all particles, so with pi_t = softmax(m_t):
 *   d loglik = sum_t sum_i (pi_t,i - 1/J) alpha delta_{t-1},i + pi_t,i dm_t,i,
 *   delta_t,j = alpha delta_{t-1},a_t(j) + dm_t,a_t(j)      (wF's tangent)
 * whose adjoint runs backwards: lambda_{t-1,i} = alpha (pi_t,i - 1/J +
 * C_t,i), C_t,i = sum over the children j of i (a_t(j) = i) of lambda_t,j, and
 * each density's weight is g_t,i = pi_t,i + C_t,i.  The state adjoint adds
 * g dm/dv and phi times the children's state adjoints.  Sums in double. */
static double sv_mop_grad_one(const float* th_est, const float* ys, int T, int64_t J, float alpha, key_t2 key,
                              float* Vt, float* Zt, float* Mt, int32_t* At, float* V0s, float* Z0s, double* lam,
                              double* lamc, double* xb, double* xbc, double* grad) {
  const float mu = th_est[0], phi = d_tanh(th_est[1]), sig = d_exp(th_est[2]), rho = d_tanh(th_est[3]);
  const float V0 = th_est[4], S0 = d_exp(th_est[5]);
  const float omr2 = 1.0f - rho * rho, k = 1.0f / omr2, c0 = -0.918938533f - 0.5f * d_log(omr2);
  key_t2 run = split_at(key, 0);
  for (int64_t j = 0; j < J; ++j) {
    Z0s[j] = normal_at(split_at(key, 1 + (uint64_t)j), 0);
    V0s[j] = V0 + S0 * Z0s[j];
  }
  const double logJ = (double)d_log((float)J);
  double ll = 0.0;
  uint64_t* C = (uint64_t*)malloc(sizeof(uint64_t) * J);
  float* m_max = (float*)malloc(sizeof(float) * T);
  for (int t = 0; t < T; ++t) {
    key_t2 sub = split_at(run, 1);
    run = split_at(run, 0);
    key_t2 sk = split_at(sub, 1);
    float* v_t = Vt + (int64_t)t * J;
    const float* v_p = t == 0 ? V0s : Vt + (int64_t)(t - 1) * J;
    const int32_t* a_p = t == 0 ? NULL : At + (int64_t)(t - 1) * J;
    float m = -INFINITY;
    for (int64_t j = 0; j < J; ++j) {
      float v = v_p[a_p ? a_p[j] : j];
      float z = normal_at(split_at(sk, (uint64_t)j), 0);
      v = mu + phi * (v - mu) + sig * z;
      float u = ys[t] * d_exp(-0.5f * v) - rho * z;
      float g = c0 - 0.5f * v - 0.5f * u * u * k;
      v_t[j] = v;
      Zt[(int64_t)t * J + j] = z;
      Mt[(int64_t)t * J + j] = g;
      if (g > m) m = g;
    }
    m_max[t] = m;
    const uint64_t total = fixed_prefix(Mt + (int64_t)t * J, m, J, C);
    ll += step_loglik(m, total, J) - logJ;
    key_t2 rk = split_at(run, 1);
    run = split_at(run, 0);
    if (t < T - 1) resample(C, total, J, uniform01(bits_at(rk, 0)), At + (int64_t)t * J);
  }
  /* reverse sweep; grad in natural parameters (mu, phi, sigma, rho, V0, S0) */
  double gn[6] = {0, 0, 0, 0, 0, 0};
  const double kd = k, rhod = rho, phid = phi, mud = mu;
  for (int64_t j = 0; j < J; ++j) {
    lam[j] = 0.0;
    xb[j] = 0.0;
  }
  for (int t = T - 1; t >= 0; --t) {
    const float* m_t = Mt + (int64_t)t * J;
    const float* v_t = Vt + (int64_t)t * J;
    const float* z_t = Zt + (int64_t)t * J;
    /* children sums through a_t of lambda_t and of the next states' adjoints */
    for (int64_t j = 0; j < J; ++j) {
      lamc[j] = 0.0;
      xbc[j] = 0.0;
    }
    if (t < T - 1) {
      const int32_t* a_t = At + (int64_t)t * J;
      for (int64_t j = 0; j < J; ++j) {
        lamc[a_t[j]] += lam[j];
        xbc[a_t[j]] += phid * xb[j];
      }
    }
    double s = 0.0;
    for (int64_t j = 0; j < J; ++j) s += exp((double)m_t[j] - (double)m_max[t]);
    const double iJ = 1.0 / (double)J;
    for (int64_t j = 0; j < J; ++j) {
      const double pi = exp((double)m_t[j] - (double)m_max[t]) / s;
      const double g = pi + lamc[j];
      /* density derivatives at (v, eta = z) */
      const double v = v_t[j], z = z_t[j];
      const double ry = (double)ys[t] * exp(-0.5 * v);
      const double u = ry - rhod * z;
      const double dm_dv = -0.5 + 0.5 * u * ry * kd;
      gn[3] += g * (rhod * kd + u * z * kd - u * u * rhod * kd * kd);
      /* state adjoint of v_t,j, then the step that made it */
      const double xbar = g * dm_dv + xbc[j];
      const double vprev = t == 0 ? (double)V0s[j] : (double)Vt[(int64_t)(t - 1) * J + At[(int64_t)(t - 1) * J + j]];
      gn[0] += xbar * (1.0 - phid);
      gn[1] += xbar * (vprev - mud);
      gn[2] += xbar * z;
      xb[j] = xbar;
      lam[j] = t > 0 ? (double)alpha * (pi - iJ + lamc[j]) : 0.0;
    }
    if (t == 0) {
      /* rinit: v_init = V0 + S0 z0, its adjoint phi times the first step's */
      for (int64_t j = 0; j < J; ++j) {
        const double xi = phid * xb[j];
        gn[4] += xi;
        gn[5] += xi * (double)Z0s[j];
      }
    }
  }
  /* chain rule to the estimation scale */
  grad[0] = gn[0];
  grad[1] = gn[1] * (1.0 - phid * phid);
  grad[2] = gn[2] * (double)sig;
  grad[3] = gn[3] * (1.0 - rhod * rhod);
  grad[4] = gn[4];
  grad[5] = gn[5] * (double)S0;
  free(C);
  free(m_max);
  return ll;
}

/* R independent SV MOP value + gradient computations, one per OpenMP thread;
 * grad (R, 6) d loglik / d theta_est. */
int cpu_sv_mop_grad(const float* theta_est, const float* ys, int T, int64_t J, float alpha, const uint32_t* keys,
                    int R, double* loglik, double* grad) {
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < R; ++r) {
    const int64_t TJ = (int64_t)T * J;
    float* Vt = (float*)malloc(sizeof(float) * TJ);
    float* Zt = (float*)malloc(sizeof(float) * TJ);
    float* Mt = (float*)malloc(sizeof(float) * TJ);
    int32_t* At = (int32_t*)malloc(sizeof(int32_t) * TJ);
    float* V0s = (float*)malloc(sizeof(float) * J);
    float* Z0s = (float*)malloc(sizeof(float) * J);
    double* w = (double*)malloc(sizeof(double) * 4 * J);
    if (!Vt || !Zt || !Mt || !At || !V0s || !Z0s || !w) {
      err = 1;
    } else {
      key_t2 k = {keys[2 * r], keys[2 * r + 1]};
      loglik[r] = sv_mop_grad_one(theta_est, ys, T, J, alpha, k, Vt, Zt, Mt, At, V0s, Z0s, w, w + J, w + 2 * J,
                                  w + 3 * J, grad + 6 * (int64_t)r);
    }
    free(Vt);
    free(Zt);
    free(Mt);
    free(At);
    free(V0s);
    free(Z0s);
    free(w);
  }
  return err;
}

int cpu_num_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}
