/*
 * mipomp.h -- C ABI of the MI355X particle-filter engine (libmipomp.so).
 *
 * Plain C: pointers, sizes and plain structs; no C++ or torch types.  Every
 * pointer argument is DEVICE memory unless its name starts with `h_`.  Every
 * compute entry point takes a HIP stream (`void*` = hipStream_t, NULL = the
 * null stream), enqueues work and returns without synchronising.  The caller
 * owns all memory; scratch comes from a caller-provided workspace whose size
 * the matching *_workspace_size() query returns.  Return value: MP_OK or an
 * error code; mp_last_error() gives the message of the calling thread's last
 * failure.  No C++ exception crosses this boundary.
 *
 * Each entry point replaces a reference function of pypomp 0.4.10.2; the
 * reference binds none of them itself (it is pure Python on JAX).  The Python
 * host layer of this package (pypomp_amd/_native.py, ctypes) is the binding.
 */
#ifndef MIPOMP_H
#define MIPOMP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MP_OK 0
#define MP_ERR_ARG 1       /* invalid argument (message in mp_last_error) */
#define MP_ERR_HIP 2       /* a HIP runtime call failed */
#define MP_ERR_WORKSPACE 3 /* workspace smaller than the size query returned */

/* models (device functors compiled into the library) */
#define MP_MODEL_LG 1   /* pypomp/models/linear_gaussian.py:139-191 */
#define MP_MODEL_SEIR 2 /* Euler-binomial SEIR, NegBin reports */
#define MP_MODEL_SV 3   /* stochastic volatility with leverage */
#define MP_MODEL_SPX 4  /* pypomp/models/spx.py:56-151: Heston-type volatility with the previous
                           return as covariate y_prev; per-particle or @vectorized rproc */

/* algorithms / forms */
#define MP_ALGO_PF 0
#define MP_ALGO_MOP 1
#define MP_ALGO_PER 2
#define MP_FORM_BATCHED 0 /* one persistent launch, one workgroup per filter */
#define MP_FORM_PLAIN 1   /* one filter at a time over the whole GPU, launches per step */

/* output flags of mp_pfilter (pypomp/core/algorithms/pfilter.py:146-156) */
#define MP_CLL 1
#define MP_ESS 2
#define MP_FILTER_MEAN 4
#define MP_PREDICTION_MEAN 8
#define MP_SHOULD_TRANS 16 /* theta on the estimation scale */

typedef struct mp_model_desc {
  int32_t kind;        /* MP_MODEL_* */
  int32_t S;           /* state words per particle */
  int32_t d;           /* parameters */
  int32_t dy;          /* observation words */
  int32_t dx;          /* LG state dimension (<= 4) */
  int32_t lg_sep_init; /* LG: initial Cholesky factor P follows X0 in theta */
  uint32_t accum_mask; /* bit s: state word s is reset to 0 at each observation
                          interval (reference: accumvars, model_mechanics.py:921) */
  int32_t identity_trans; /* 1: the estimation scale is the natural scale (the
                          reference's ParTrans() with no transforms,
                          core/par_trans.py:70-111); 0: the model's own transforms */
  int32_t vectorized;  /* 1: the model's rproc is @vectorized (model_mechanics.py:
                          311-355, 553-568): each Euler sub-step draws particle j's
                          variates at flat index j of one draw from the sub-step key,
                          instead of from split(key, J)[j]; models that define it only */
} mp_model_desc;

/* Observation series of U units (U = 1 for a single Pomp).  Unit u owns rows
 * [obs_off[u], obs_off[u+1]) of ys/times/nstep/tidx0; its Euler step sizes
 * start at dt_ext + dt_off[u]; tidx0 = unit-local index of the first Euler
 * step of each observation interval (reference: _calc_ys_covars,
 * core/algorithms/helpers.py:279-299). */
typedef struct mp_series_desc {
  const float* ys;        /* (sum T_u, dy) */
  const float* times;     /* (sum T_u) */
  const int32_t* nstep;   /* (sum T_u) */
  const int32_t* tidx0;   /* (sum T_u) */
  const float* dt_ext;    /* (sum_u sum_t nstep) */
  const int64_t* obs_off; /* (U + 1) */
  const int64_t* dt_off;  /* (U) */
  float t0;
  int32_t U;
  int32_t Tmax;           /* max_u T_u */
  int64_t max_euler;      /* max_u sum_t nstep */
  /* Covariates interpolated onto each unit's Euler grid (reference:
   * covars_extended from _calc_ys_covars / _calc_interp_covars,
   * core/algorithms/helpers.py:248-299): unit u owns E_u + 1 rows from
   * cv_off[u], E_u = its number of Euler steps; row tidx0[t] + k holds the
   * covariates of sub-step k of interval t (model_mechanics.py:894-906), row
   * tidx0[t] + nstep[t] those of observation t (pfilter.py:124-128), row 0 those
   * of rinit (pfilter.py:28-35).  NULL with dc = 0. */
  const float* covars;    /* (sum_u (E_u + 1), dc) */
  const int64_t* cv_off;  /* (U) */
  int32_t dc;             /* covariates per row */
  int32_t max_nstep;      /* max over intervals of nstep (the reference's
                             _max_steps_per_interval, helpers.py:296) */
} mp_series_desc;

const char* mp_last_error(void);
int32_t mp_abi_version(void);

/* ---- host-side key utilities (CPU; no device needed) ---------------------
 * Threefry-2x32-20 and jax.random.split in partitionable mode, used by the host
 * layer for key derivation (reference: Pomp._update_fresh_key,
 * core/pomp.py:391-408; rep keys, core/estimation_mixin.py:188-190). */
int32_t mp_host_threefry2x32(const uint32_t* h_keys, const uint32_t* h_ctr, uint32_t* h_out, int64_t n);
int32_t mp_host_split(const uint32_t* h_key, int64_t n, uint32_t* h_out);

/* ---- device primitives (parity tests and building blocks) ----------------
 * keys: (nkeys, 2) uint32.  For each key k and flat index i < n:
 *   bits[k, i]    = random_bits(k, 32, (n,))[i]     (jax.random.bits)
 *   uniform[k, i] = jax.random.uniform(k, (n,))[i]  (helpers.py:28 call site)
 *   normal[k, i]  = jax.random.normal(k, (n,))[i]   (linear_gaussian.py:147,162)
 *   split[k, i]   = jax.random.split(k, n)[i]       (pfilter.py:23,105,160) */
int32_t mp_random_bits(const uint32_t* keys, int64_t nkeys, int64_t n, uint32_t* out, void* stream);
int32_t mp_uniform(const uint32_t* keys, int64_t nkeys, int64_t n, float* out, void* stream);
int32_t mp_normal(const uint32_t* keys, int64_t nkeys, int64_t n, float* out, void* stream);
int32_t mp_split(const uint32_t* keys, int64_t nkeys, int64_t n, uint32_t* out, void* stream);

/* Systematic resampling of R independent rows of J log-weights (row r uses
 * key r: u = uniform(key), comb (u + j) / J), replacing _resample
 * (core/algorithms/helpers.py:14-31) with its order-independent fixed-point
 * definition; ancestors (R, J) int32, already clamped to J - 1 like the
 * reference's gather (helpers.py:87). */
size_t mp_resample_workspace_size(int64_t R, int64_t J);
int32_t mp_resample_systematic(const float* logw, const uint32_t* keys, int64_t R, int64_t J,
                               int32_t* ancestors, void* workspace, size_t ws_bytes, void* stream);

/* ---- filters ---------------------------------------------------------------
 * R filters; filter f uses key keys[f], unit unit[f] (NULL: unit 0) and
 * parameters theta + f * theta_stride (theta_stride 0: shared).  loglik (R)
 * float64 sums over t of the per-observation float32-evaluated terms. */
size_t mp_filter_workspace_size(int32_t algo, const mp_model_desc* h_model, const mp_series_desc* h_series,
                                int32_t R, int32_t J, int32_t form);

/* Bootstrap particle filter: _pfilter_internal (core/algorithms/pfilter.py:15-182),
 * batched over replicates/parameter sets like _vmapped_pfilter_internal2
 * (pfilter.py:185-207).  cll/ess (R, Tmax), fmean/pmean (R, Tmax, S) may be
 * NULL unless the matching flag is set. */
int32_t mp_pfilter(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                   int64_t theta_stride, const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J,
                   float thresh, int32_t flags, int32_t form, double* loglik, float* cll, float* ess,
                   float* fmean, float* pmean, void* workspace, size_t ws_bytes, void* stream);

/* mp_pfilter without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * results and errors as mp_pfilter; its scratch is a fixed array of
 * MP_NOALLOC_SCRATCH_BYTES in the library, one per GPU, held until the process
 * ends.  A call whose workspace size (mp_filter_workspace_size) exceeds it
 * returns MP_ERR_ARG before launching anything.  The call allocates nothing
 * and does not record kernel timings.  Every call on one GPU uses that GPU's
 * one array: calls that overlap on a GPU (on different streams or threads)
 * race and must be serialised by the caller. */
#define MP_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int32_t mp_pfilter_noalloc(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                           int64_t theta_stride, const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J,
                           float thresh, int32_t flags, int32_t form, double* loglik, float* cll, float* ess,
                           float* fmean, float* pmean, void* stream);

/* MOP(alpha) log-likelihood value: _mop_internal (core/algorithms/mop.py:19-135);
 * theta on the estimation scale.  Returns +loglik (the reference's F.mop
 * returns its negative, functional/mop.py:58-62). */
int32_t mp_mop(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
               int64_t theta_stride, const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J,
               float alpha, int32_t form, double* loglik, void* workspace, size_t ws_bytes, void* stream);

/* MOP(alpha) value and its gradient d loglik / d theta_est by a hand-written
 * adjoint: the reverse-mode derivative of _mop_internal (mop.py:19-135) that
 * jax.grad computes through _mop_internal_mean (ad_helpers.py:6-8; the
 * reference's gradient is -grad / T), ancestry and random numbers held fixed.
 * The forward keeps a checkpoint every K observations; the reverse sweep
 * recomputes each K-step block from it.  One Euler step per observation.
 * unit (R) the panel unit of each filter (its own rows and length; the
 * panel train, core/algorithms/train.py:191-415, differentiates one MOP per
 * unit), or NULL for unit 0 of a one-unit series.  loglik (R) float64, grad
 * (R, d) float32.  form MP_FORM_PLAIN runs the
 * filters one launch at a time.  Models with an adjoint: SV, and LG with
 * dx = dy = 1 or dx = dy = 2; others return MP_ERR_ARG.  The train loop
 * (core/algorithms/train.py:38-180) calls this once per gradient step. */
size_t mp_mop_grad_workspace_size(const mp_model_desc* h_model, const mp_series_desc* h_series, int32_t R,
                                  int32_t J, int32_t K);
int32_t mp_mop_grad(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                    int64_t theta_stride, const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J,
                    float alpha, int32_t K, int32_t form, double* loglik, float* grad, void* workspace,
                    size_t ws_bytes, void* stream);

/* Hessian of the MOP(alpha) log-likelihood with respect to theta_est: the
 * second derivative jax.hessian takes of _mop_internal_mean
 * (core/algorithms/ad_helpers.py:8, 56-75; the reference's Hessian is
 * -hess / T), ancestry and random numbers held fixed, the stop-gradient
 * weights held at theta (mop.py:125).  Forward mode over hyper-dual numbers,
 * one workgroup per (row, parameter pair); the key tree and resampling of
 * mp_mop, whose value loglik (R) float64 reproduces.  grad (R, d) and hess
 * (R, d, d) float32, hess symmetric.  U = 1.  Models: SV, and LG with
 * dx = dy = 1 or 2; others return MP_ERR_ARG.  Newton and WeightedNewton
 * (core/optimizer.py:270-330) of the train loop call it once per step. */
size_t mp_mop_hess_workspace_size(const mp_model_desc* h_model, const mp_series_desc* h_series, int32_t R,
                                  int32_t J);
int32_t mp_mop_hess(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                    int64_t theta_stride, const uint32_t* keys, int32_t R, int32_t J, float alpha, double* loglik,
                    float* grad, float* hess, void* workspace, size_t ws_bytes, void* stream);

/* The model's estimation-to-natural transform (from_est) applied on the host
 * to n rows of d parameters in place, with the device's arithmetic (the same
 * deterministic exp/tanh/logistic), so host-side traces and priors see the
 * values the kernels see.  The reference applies ParTrans.from_est to
 * parameter arrays on the host (core/par_trans.py:70-111).  No-op for
 * identity_trans. */
int32_t mp_host_from_est(const mp_model_desc* h_model, float* h_theta, int64_t n);

/* Simulation of states and observations: _simulate_internal /
 * _simulate_helper (core/algorithms/simulate.py:15-142) vmapped over R
 * parameter rows (simulate.py:71-74; functional/simulate.py:11-78), with the
 * reference's key tree and its time bookkeeping (the process of interval i
 * starts at t0 for i <= 1 and at times[i-2] after; rmeas sees times0[i]).
 * theta (R, d) natural scale unless flags has MP_SHOULD_TRANS; keys (R, 2);
 * times (T) or NULL for the series' own; U = 1.  X (R, T+1, S, nsim) and
 * Y (R, T, dy, nsim) float32: the reference's internal layout before its
 * transpose to (R, nsim, T+1, S).  No workspace. */
int32_t mp_simulate(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* theta,
                    int64_t theta_stride, const uint32_t* keys, const float* times, int32_t R, int32_t nsim,
                    int32_t flags, float* X, float* Y, void* stream);

/* One IF2 iteration m: _perfilter_internal (core/algorithms/mif.py:88-258) for R
 * swarms.  thetas_in/thetas_out (R, J, d) on the estimation scale; sigmas and
 * sigmas_init (d) as RWSigma.sigmas_array / sigmas_init_array
 * (core/rw_sigma.py:184-191).  Cooling (RWSigma.cooling_factor,
 * rw_sigma.py:193-225): cooling NULL takes the geometric schedule
 * cool^(nt/T_u + m) with cool = a^(1/50), evaluated in-kernel; otherwise
 * cooling (sum T_u) holds this iteration's factors c(nt, m, T_u) at the
 * observation rows of the series (cosine, hyperbolic and custom schedules,
 * evaluated by the caller).  theta_mean (R, d): mean of thetas_out over J.
 * ancestry (R, J) int32 or NULL: the composed resampling ancestry of the final
 * particles (mif.py:237-245, return_ancestry; panel IF2 with block=False,
 * core/algorithms/panel_mif.py:224-231).  Panel units of one call (unit[f])
 * are independent filters; the panel driver (panel_mif.py:17-247) runs units
 * that share parameters one call after another.  Filters are taken longest
 * unit first from a device work queue, so uneven units balance over the CUs. */
int32_t mp_perfilter(const mp_model_desc* h_model, const mp_series_desc* h_series, const float* thetas_in,
                     const uint32_t* keys, const int32_t* unit, int32_t R, int32_t J, float thresh,
                     const float* sigmas, const float* sigmas_init, float cool, const float* cooling, int32_t m,
                     float* thetas_out, float* theta_mean, double* loglik, int32_t* ancestry, void* workspace,
                     size_t ws_bytes, void* stream);

/* ---- launch-shape tuning ------------------------------------------------------
 * No counterpart in the reference (XLA picks its own launch shapes).  Process-
 * wide settings read when a call is planned (workspace sizes depend on them:
 * query the size after setting).  The Python autotuner
 * (pypomp_amd/autotune.py) sweeps them on the caller's workload and caches the
 * best values. */
/* The four plain-form shapes take 0 (the default) for automatic: scan 1024 x 4 and propagate
   512 x 2 with one or two lanes, scan 256 x 16 and propagate 256 x 4 with three or four. */
#define MP_TUNE_PLAIN_SCAN_ROWS 1 /* consecutive particles per scan thread: 1, 2, 4, 8, 16 */
#define MP_TUNE_PLAIN_PROP_ROWS 2 /* particle rows per propagate block, 1..16 */
#define MP_TUNE_BATCHED_BLOCK 3   /* batched workgroup size: 0 automatic, 256, 512, 768, 1024 (SEIR pfilter and MOP only; 768 elsewhere) */
#define MP_TUNE_GRAD_BLOCK 4      /* MOP-adjoint workgroup size: 0 automatic, 256, 512 */
#define MP_TUNE_PLAIN_SCAN_THREADS 5 /* plain scan workgroup size: 256, 512, 1024 */
#define MP_TUNE_PLAIN_PROP_THREADS 6 /* plain propagate workgroup size: 256, 512, 1024 */
#define MP_TUNE_PLAIN_LDS_TILES 7    /* scan tiles whose ends each propagate block forms in LDS,
                                        0..1024 (default 1024); more tiles (or 0) take a separate
                                        launch per step that writes them to global memory */
#define MP_TUNE_PLAIN_XCD_MAP 8      /* 1 (default): propagate blocks take the particle chunks
                                        of the scan tiles on their own XCD; 0: block order */
#define MP_TUNE_PLAIN_LANES 9        /* replicate filters the plain form runs at once, each on
                                        its own stream (default priority) with its own buffers:
                                        1..4 (default 4; the no-allocation entry point takes 1) */
int32_t mp_tune_set(int32_t param, int32_t value);
int32_t mp_tune_get(int32_t param);

/* ---- kernel timing ----------------------------------------------------------
 * No counterpart in the reference (it has no profiler hooks).  While enabled,
 * every kernel launch of the entry points above is bracketed by HIP events on
 * its stream; mp_profile_read synchronises on them and returns, per kind
 * (0 plain propagate, 1 plain scan, 2 batched filter, 3 key schedule,
 * 4 VALU probe), the
 * summed milliseconds and launch counts, then clears the records. */
int32_t mp_profile_enable(int32_t on);

/* VALU issue calibration, no counterpart in the reference: blocks x 256
 * threads each run iters rounds of eight independent dependency chains of
 * one instruction mix (kind 0: float FMAs; 1: threefry's integer add, rotate
 * and xor; 2: FMAs with base-2 exponentials; 3: FMAs with integer add and xor;
 * 4: integer add and xor alone), writing one float per thread to
 * out (blocks x 256).  Timed as kind 4 of mp_profile_read; with the VALU
 * instruction counters it gives the issue ceiling the kernels are held to. */
int32_t mp_valu_probe(int32_t kind, int64_t iters, int32_t blocks, float* out, void* stream);
int32_t mp_profile_read(double* h_ms, int64_t* h_count, int32_t nkinds);

#ifdef __cplusplus
}
#endif

#endif /* MIPOMP_H */
