// Internal launch descriptions shared by the kernel translation units and the
// C-ABI layer.  Plain data only; every pointer is device memory unless noted.
#pragma once

#include "models.hpp"
#include <hip/hip_ext.h>

namespace mp {

// Kernel timing for mp_profile_enable: the launch wrapper of the C ABI sets
// these events, and the next launch attaches them to its own dispatch packet
// (start and end timestamps of that kernel alone, with no event packets of
// their own in the stream), then clears them.
struct LaunchTiming {
  hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchTiming g_launch_timing;

template <typename K, typename... Args>
inline void mp_launch(K kernel, dim3 grid, dim3 block, uint32_t lds, hipStream_t s, Args... args) {
  if (g_launch_timing.start) {
    const LaunchTiming lt = g_launch_timing;
    g_launch_timing = LaunchTiming{};
    hipExtLaunchKernelGGL(kernel, grid, block, lds, s, lt.start, lt.stop, 0u, args...);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
  }
}


enum Algo : int { ALGO_PF = 0, ALGO_MOP = 1, ALGO_PER = 2 };

// A set of observation series (one per panel unit; U = 1 for a single Pomp).
// Unit u owns rows [obs_off[u], obs_off[u+1]) of ys/times/nstep/tidx0 and the
// Euler step sizes starting at dt_ext + dt_off[u]; tidx0 is the unit-local
// index of the first Euler step of each observation interval.
struct SeriesSet {
  const float* ys;       // (sum T_u, dy)
  const float* times;    // (sum T_u)
  const int32_t* nstep;  // (sum T_u)
  const int32_t* tidx0;  // (sum T_u)
  const float* dt_ext;   // (sum_u sum nstep)
  const int64_t* obs_off;  // (U + 1)
  const int64_t* dt_off;   // (U)
  float t0;
  int U;
  int Tmax;
  int64_t max_euler;  // max over units of the number of Euler steps
  // Covariates on each unit's Euler grid: unit u's rows start at cv_off[u];
  // row tidx0[t] + k is the k-th sub-step of interval t and row
  // tidx0[t] + nstep[t] the observation after it (the reference's
  // covars_extended, one row per grid time plus the last observation time).
  const float* covars;     // (sum_u (E_u + 1), dc), or null when dc == 0
  const int64_t* cv_off;   // (U), or null when dc == 0
  int dc;                  // covariates per row
  int max_nstep;           // max over intervals of nstep
};

// Row of covariates at grid index i of unit u (null without covariates).
__device__ __forceinline__ const float* cov_row(const SeriesSet& s, int u, int64_t i) {
  return s.dc ? s.covars + (s.cv_off[u] + i) * s.dc : nullptr;
}

// Key-schedule words (keys) per filter: PF/MOP sub-step keys + one resampling
// key per observation; IF2 adds a perturbation key per observation and two
// initial keys.
__host__ __device__ inline int64_t sched_stride(int algo, const SeriesSet& s) {
  return algo == ALGO_PER ? 2 + s.max_euler + 2 * (int64_t)s.Tmax : s.max_euler + (int64_t)s.Tmax;
}

enum Flags : int {
  F_CLL = 1,
  F_ESS = 2,
  F_FMEAN = 4,
  F_PMEAN = 8,
  F_SHOULD_TRANS = 16,  // theta is on the estimation scale (apply from_est)
};

struct BatchArgs {
  ModelDesc md;
  SeriesSet ser;
  const int32_t* unit;      // (R) unit of each filter, or null (unit 0)
  const float* theta;       // PF/MOP: (R, d) or shared (stride 0); PER: (R, J, d) est. scale
  int64_t theta_stride;     // floats between filters' thetas (PF/MOP)
  const uint32_t* keys;     // (R, 2) filter keys
  const uint32_t* sched;    // key schedules, filter f at sched + 2 * f * sched_len
  int64_t sched_len;        // keys per filter schedule
  int R, J;
  int flags;
  float logthresh;          // log(thresh); -inf when thresh == 0
  float alpha;              // MOP
  // IF2
  const float* sigmas;      // (d) non-initial sigmas
  const float* sigmas_init; // (d)
  float cool_factor;        // geometric cooling factor a^(1/50) (float32 of the double)
  const float* cool_table;  // or the factors c(nt, m) of this iteration per observation row (other schedules)
  int mif_m;                // iteration index m
  float* theta_out;         // (R, J, d) swarm after the final resample (est. scale)
  float* theta_mean;        // (R, d) mean of theta_out
  int32_t* ancestry;        // (R, J) composed final ancestry, or null
  // outputs
  double* loglik;           // (R)
  float* cll;               // (R, Tmax)
  float* ess;               // (R, Tmax)
  float* fmean;             // (R, Tmax, S)
  float* pmean;             // (R, Tmax, S)
  // per-filter scratch (global) when not LDS-resident
  char* ws;
  int64_t ws_stride;
  int lds_resident;
  // work queue: filters are taken in order[] (longest unit first) via *queue
  const int32_t* order;
  int32_t* queue;
};

// Per-filter scratch layout (bytes), shared by host sizing and device use.
struct FilterScratch {
  int64_t off_x0, off_x1, off_lw, off_c, total;
};

__host__ __device__ inline int64_t align256(int64_t v) { return (v + 255) & ~int64_t(255); }

__host__ __device__ inline FilterScratch filter_scratch_layout(int W, int J) {
  FilterScratch s;
  int64_t xs = (int64_t)W * J * 4;
  s.off_x0 = 0;
  s.off_x1 = align256(xs);
  s.off_lw = s.off_x1 + align256(xs);
  s.off_c = s.off_lw + align256((int64_t)J * 4);
  s.total = s.off_c + align256((int64_t)J * 8);
  return s;
}

// Scalars of observation t of a plain-form filter, one row per t, built by
// k_plain_reset before the first step, so that a kernel starts with one
// independent load instead of a chain of dependent ones.
struct PlainStep {
  int64_t obs;         // row of ys / times: obs_off[unit] + t
  int64_t base;        // schedule index of the first sub-step key: tidx0 + t
  int ti0;             // first Euler sub-step (index into the unit's dt_ext)
  int nst;             // Euler sub-steps
  float tstart, ty;    // start of the interval, observation time
  float dt0;           // first sub-step length (0 without sub-steps)
  float ures_prev;     // comb uniform of step t - 1's resampling key (0 at t = 0)
  uint32_t k0, k1;     // first sub-step key
};

// Running state of a plain-form filter: the FP64 sum of the log-likelihood
// terms of the steps finished so far, kept by block 0 of each propagate launch.
struct PlainState {
  double loglik;
};

// Words of one block's diagnostic sums: ESS, filter mean and prediction mean
// of up to 7 state words.
constexpr int kDiagWords = 17;  // 1 + 2 S words for S <= 8

// XCDs of an MI355X: the hardware dispatches workgroups over them round-robin
// by block index, and each has its own L2.
constexpr int kXcds = 8;

struct PlainArgs {
  ModelDesc md;
  SeriesSet ser;
  int unit;
  const float* theta;   // (d)
  const void* params;   // the model's prepared parameters (M::P) for theta
  const uint32_t* fkey; // (2) filter key
  const uint32_t* sched;
  int J, t, T, flags, f;
  float logthresh;
  float* X0;
  float* X1;
  float* lw;
  uint64_t* C;          // (J) prefix sums of the fixed-point weights within each scan tile
  uint64_t* agg;        // (ntiles) sum of each scan tile's fixed-point weights
  uint64_t* ends;       // (ntiles) inclusive prefix of agg (k_plain_ends), or null: every
                        // propagate block forms it in LDS from agg
  float* pmax;          // (nprop) per propagate block max / min log-weight
  float* pmin;
  float* mm;            // (2, 2) per step parity: max and min log-weight of the step
  PlainStep* steps;     // (T) step scalars
  PlainState* st;
  double* loglik_out;   // (1) written after the last step
  float* cll;           // (T) or null
  float* ess;           // (T) or null
  float* fmean;         // (T, S) or null
  float* pmean;         // (T, S) or null
  double* diag_part;    // (ndiag, kDiagWords) per-block diagnostic sums
  int ndiag;            // diagnostic blocks
  int rows;             // particles per scan thread (tile = rows * scan workgroup size)
  int tile_len;         // particles per scan tile (a power of two)
  int tile_shift;       // log2(tile_len)
  int ntiles;           // scan tiles
  int rows_prop;        // particle rows per propagate block
  int prop_len;         // particles per propagate block (<= rows_prop * blockDim)
  int nprop;            // propagate blocks
  int xcd_q;            // > 0: propagate block b takes chunk xcd_chunk(b, xcd_q); 0: chunk b
  int win;              // LDS ancestor-window entries of a propagate block
  float logJ, log1J;    // d_log(J) and d_log(1 / J) (float32)
  int sh;               // fixed-point weight shift, shift_for(J)
};

// At most this many scan tiles have their ends formed in LDS by each
// propagate block; more take a k_plain_ends launch per step.
constexpr int kMaxLdsTiles = 1024;

// MOP value + adjoint gradient launch (one workgroup per filter).
struct GradArgs {
  ModelDesc md;
  SeriesSet ser;
  const int32_t* unit; // (R) panel unit of each filter, or null (unit 0)
  const float* theta;  // (R, d) estimation scale
  int64_t theta_stride;
  const uint32_t* keys;
  const uint32_t* sched;
  int64_t sched_len;
  int R, J, K;
  float alpha;
  double* loglik;  // (R)
  float* grad;     // (R, d): d loglik / d theta_est
  char* ws;
  int64_t ws_stride;
  int lds_work;    // per-step working arrays in LDS (else in the global scratch)
};

// MOP Hessian launch (one workgroup per (parameter row, parameter pair)).
struct HessArgs {
  ModelDesc md;
  SeriesSet ser;
  const float* theta;  // (R, d) estimation scale
  int64_t theta_stride;
  const uint32_t* keys;
  const uint32_t* sched;
  int64_t sched_len;
  int R, J;
  float alpha;
  double* loglik;  // (R)
  float* grad;     // (R, d): d loglik / d theta_est
  float* hess;     // (R, d, d): d^2 loglik / d theta_est^2
  char* ws;        // per resident workgroup: hyper-dual particle buffers
  int64_t ws_stride;
};

// Scratch of one Hessian workgroup: states and weights as hyper-dual numbers
// (32 bytes each), ping-pong, plus the log-weights and prefix sums.
struct HessScratch {
  int64_t off_x0, off_x1, off_w0, off_w1, off_lw, off_c, total;
};
constexpr int64_t kHdBytes = 32;
__host__ __device__ inline HessScratch hess_scratch_layout(int S, int J) {
  HessScratch s;
  const int64_t xs = align256((int64_t)S * J * kHdBytes), ws = align256((int64_t)J * kHdBytes);
  s.off_x0 = 0;
  s.off_x1 = xs;
  s.off_w0 = 2 * xs;
  s.off_w1 = 2 * xs + ws;
  s.off_lw = 2 * xs + 2 * ws;
  s.off_c = s.off_lw + align256((int64_t)J * 4);
  s.total = s.off_c + align256((int64_t)J * 8);
  return s;
}

// Simulation launch (one thread per (row, simulation)); U = 1.
constexpr int kSimMaxDy = 8;
// bytes per filter of prepared model parameters in the plain form (the LG
// model of dimension 8 takes 1584)
constexpr int64_t kParamSlot = 2048;
struct SimArgs {
  ModelDesc md;
  SeriesSet ser;
  const float* theta;   // (R, d), natural scale unless F_SHOULD_TRANS
  int64_t theta_stride;
  const uint32_t* keys; // (R, 2)
  const float* times;   // (T) observation times (the series' own or the caller's)
  int R, nsim;
  int flags;
  float* X;             // (R, T+1, S, nsim)
  float* Y;             // (R, T, dy, nsim)
};

}  // namespace mp
