// Batched ("fused") form: one workgroup runs a whole filter -- every time step,
// every Euler sub-step, the measurement density, the log-sum-exp and the
// systematic resampling -- in one persistent launch; the grid strides over
// filters (replicates x parameter sets x panel units).  Per-filter particle
// buffers live in LDS when they fit, otherwise in an L2-resident global slice.
//
// Algorithms (template ALGO):
//  ALGO_PF  -- bootstrap filter: _pfilter_internal/_pfilter_step
//              (pypomp/core/algorithms/pfilter.py:15-182)
//  ALGO_MOP -- MOP(alpha) value: _mop_internal/_mop_step (mop.py:19-135)
//  ALGO_PER -- IF2 perturbed filter: _perfilter_internal/_perfilter_scan_body
//              (mif.py:88-258), per-particle parameters carried as extra
//              state words and gathered with the particles.
#include "blockops.hpp"
#include "engine.hpp"
#include "resample.hpp"
#include "mech.hpp"

namespace mp {

__device__ __forceinline__ Key ld_key(const uint32_t* p, int64_t i) { return Key{p[2 * i], p[2 * i + 1]}; }
__device__ __forceinline__ void st_key(uint32_t* p, int64_t i, Key k) {
  p[2 * i] = k.k0;
  p[2 * i + 1] = k.k1;
}

__device__ __forceinline__ Key shfl_key(Key k, int lane) {
  return Key{(uint32_t)__shfl((int)k.k0, lane, 64), (uint32_t)__shfl((int)k.k1, lane, 64)};
}

// Key schedule of a bootstrap / MOP filter (pfilter.py:23-25,105,160 and the
// sub-step chain of model_mechanics.py:904): per observation t, at
// tidx0[t] + t, the nstep[t] sub-step keys followed by the resampling key.
// Per observation the reference's tree is (key1, sub) = split(key), the
// sub-step chain (out, sub) <- split(sub), then (key, out) <- split(key1).
// Four lanes per filter evaluate sibling splits together (one threefry
// instruction stream, a different counter per lane): lanes 0 and 1 the
// sub-step chain, lanes 2 and 3 the split of key1 alongside its first link, so
// an observation costs 1 + nstep dependent hashes instead of 2 nstep + 4.
__global__ void k_sched_pf(const uint32_t* __restrict__ keys, SeriesSet ser, const int32_t* __restrict__ unit,
                           int64_t sched_len, uint32_t* __restrict__ sched, int R) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = g >> 2, c = g & 1, hi = (g >> 1) & 1;
  if (f >= R) return;  // whole quads: all four lanes share f
  const int qb = (threadIdx.x & 63) & ~3;  // the quad's first lane in its wave
  const int u = unit ? unit[f] : 0;
  const int64_t o0 = ser.obs_off[u];
  const int T = (int)(ser.obs_off[u + 1] - o0);
  uint32_t* out = sched + 2 * (int64_t)f * sched_len;
  Key key = split_at(ld_key(keys, f), 0);  // split(key, J+1)[0]
  int64_t pos = 0;
  for (int t = 0; t < T; ++t) {
    const Key a = split_at(key, (uint64_t)c);  // lane c: split(key)[c]
    const Key key1 = shfl_key(a, qb), sub0 = shfl_key(a, qb + 1);
    const int n = ser.nstep[o0 + t];
    Key sub = sub0, next = key1;
    for (int k = 0; k < (n > 0 ? n : 1); ++k) {
      const Key r = split_at(hi ? key1 : sub, (uint64_t)c);
      if (!hi && c == 1 && k < n) st_key(out, pos + k, r);  // sub-step key k
      if (hi && c == 1 && k == 0) st_key(out, pos + n, r);  // resampling key
      const Key r0 = shfl_key(r, qb), r2 = shfl_key(r, qb + 2);
      sub = r0;
      if (k == 0) next = r2;
    }
    pos += n + 1;
    key = next;
  }
}

// Key schedule of one perturbed filter (mif.py:113-181): [init-perturbation
// key, rinit base key], then per observation t at 2 + tidx0[t] + 2t:
// [perturbation key, nstep[t] sub-step keys, resampling key].  Observation
// steps are independent (split(key, T+1)), so one thread per (filter, step).
__global__ void k_sched_per(const uint32_t* __restrict__ keys, SeriesSet ser, const int32_t* __restrict__ unit,
                            int64_t sched_len, uint32_t* __restrict__ sched, int R) {
  const int f = blockIdx.y;
  if (f >= R) return;
  const int u = unit ? unit[f] : 0;
  const int64_t o0 = ser.obs_off[u];
  const int T = (int)(ser.obs_off[u + 1] - o0);
  uint32_t* out = sched + 2 * (int64_t)f * sched_len;
  const Key k = ld_key(keys, f);
  const Key key1 = split_at(k, 0);
  const Key key2 = split_at(key1, 0);
  int t = blockIdx.x * blockDim.x + threadIdx.x - 1;
  if (t == -1) {
    st_key(out, 0, split_at(k, 1));
    st_key(out, 1, key1);
    return;
  }
  if (t >= T) return;
  const Key sk = split_at(key2, 1 + (uint64_t)t);
  int64_t pos = 2 + ser.tidx0[o0 + t] + 2 * (int64_t)t;
  st_key(out, pos++, split_at(sk, 0));
  Key proc = split_at(sk, 1);
  const int n = ser.nstep[o0 + t];
  for (int j = 0; j < n; ++j) {
    st_key(out, pos++, split_at(proc, 1));
    proc = split_at(proc, 0);
  }
  st_key(out, pos, split_at(sk, 2));
}

// Per-particle words carried by a perturbed filter: state, parameters and
// one ancestry word (the particle's original index, bit-cast to float).
template <class M>
struct WMax {
  static constexpr int v = M::MAXS + M::MAXP + 1;
};

// Sub-steps per observation whose particle-independent values fit the LDS table.
constexpr int kMaxSub = 64;

// LDS table entries: MODE 0 none, 1 the particle-independent sub-step and
// observation values (one parameter set for all particles), 2 their parts
// that depend on the time or the data only (per-particle parameters).
template <class M, int MODE>
struct HoistTables {
  struct Sub {};
  struct Obs {};
};
template <class M>
struct HoistTables<M, 1> {
  using Sub = typename M::Sub;
  using Obs = typename M::Obs;
};
template <class M>
struct HoistTables<M, 2> {
  using Sub = typename M::Tab;
  using Obs = typename M::TabObs;
};

template <class M, int ALGO>
__device__ __forceinline__ void run_filter(const BatchArgs& a, const int f, char* buf, double* red, uint64_t* tab) {
  constexpr int WM = (ALGO == ALGO_PER) ? WMax<M>::v : M::MAXS;
  const ModelDesc md = a.md;
  const int J = a.J, S = M::FIXED_S ? M::MAXS : md.S, d = md.d, dy = md.dy;
  const bool track = (ALGO == ALGO_PER) && a.ancestry != nullptr;
  const int W = S + (ALGO == ALGO_PER ? d + (track ? 1 : 0) : 0);
  const int tid = threadIdx.x, B = blockDim.x;
  const int u = a.unit ? a.unit[f] : 0;
  const int64_t o0 = a.ser.obs_off[u];
  const int T = (int)(a.ser.obs_off[u + 1] - o0);
  const float* ys = a.ser.ys + o0 * dy;
  const float* times = a.ser.times + o0;
  const int32_t* nstep = a.ser.nstep + o0;
  const int32_t* tidx0 = a.ser.tidx0 + o0;
  const float* dts = a.ser.dt_ext + a.ser.dt_off[u];
  const uint32_t* sch = a.sched + 2 * (int64_t)f * a.sched_len;
  const FilterScratch L = filter_scratch_layout(W, J);
  float* X[2] = {reinterpret_cast<float*>(buf + L.off_x0), reinterpret_cast<float*>(buf + L.off_x1)};
  float* lw = reinterpret_cast<float*>(buf + L.off_lw);
  uint64_t* Cc = reinterpret_cast<uint64_t*>(buf + L.off_c);
  const bool trans = (ALGO == ALGO_MOP) || (ALGO == ALGO_PER) || (a.flags & F_SHOULD_TRANS);
  const float fJ = (float)J;

  typename M::P P;
  if (ALGO != ALGO_PER) {
    float th[M::MAXP];
    MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < d) th[p] = a.theta[f * a.theta_stride + p];
    if (trans && !md.identity_trans) M::from_est(md, th);
    M::prepare(md, th, P);
  }

  // ---- initial particles (pfilter.py:23-35; mif.py:113-133) ----
  {
    Key rbase, isub;
    float c0 = 0.f;
    if (ALGO == ALGO_PER) {
      isub = ld_key(sch, 0);
      rbase = ld_key(sch, 1);
      if (a.cool_table) {
        c0 = a.cool_table[o0];
      } else {
        float fr = 0.0f / (float)T + (float)a.mif_m;
        c0 = d_pow(a.cool_factor, fr);
      }
    } else {
      rbase = ld_key(a.keys, f);
    }
    for (int j = tid; j < J; j += B) {
      float x[WM];
      const Key pk = split_at(rbase, 1 + (uint64_t)j);
      if (ALGO == ALGO_PER) {
        float th[M::MAXP];
        const float* tin = a.theta + ((int64_t)f * J + j) * d;
        MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < d) {
          float z = normal_at(isub, (uint64_t)j * d + p);
          x[S + p] = tin[p] + (c0 * a.sigmas_init[p]) * z;
          th[p] = x[S + p];
        }
        if (track) x[S + d] = __int_as_float(j);
        if (!md.identity_trans) M::from_est(md, th);
        typename M::P Pj;
        M::prepare(md, th, Pj);
        m_rinit<M>(md, Pj, pk, a.ser.t0, x, cov_row(a.ser, u, 0));
      } else {
        m_rinit<M>(md, P, pk, a.ser.t0, x, cov_row(a.ser, u, 0));
      }
      MP_UNROLL for (int s = 0; s < WM; ++s) if (s < W) X[0][(int64_t)s * J + j] = x[s];
    }
  }
  __syncthreads();

  int cur = 0;
  bool have_anc = false, uniform_w = true;
  float wconst = (ALGO == ALGO_PER) ? -d_log(fJ) : d_log(1.0f / fJ);
  float mop_c = d_log(1.0f / fJ);
  uint64_t total = 0;
  float ures = 0.f;
  double loglik = 0.0;
  const float logJ = d_log(fJ);
  const float nan_sub = d_log(1e-18f);

  for (int t = 0; t < T; ++t) {
    const int nst = nstep[t];
    const int ti0 = tidx0[t];
    const int64_t base = (ALGO == ALGO_PER) ? (2 + ti0 + 2 * (int64_t)t + 1) : (ti0 + (int64_t)t);
    const float ty = times[t];
    const float tstart = t == 0 ? a.ser.t0 : times[t - 1];
    const float* y = ys + (int64_t)t * dy;
    const float mop_wP = a.alpha * mop_c;
    float cft = 0.f;
    Key kpert{0, 0};
    if (ALGO == ALGO_PER) {
      if (a.cool_table) {
        cft = a.cool_table[o0 + t];
      } else {
        float fr = (float)t / (float)T + (float)a.mif_m;
        cft = d_pow(a.cool_factor, fr);
      }
      kpert = ld_key(sch, base - 1);
    }
    float lmax = -__builtin_inff(), lmin = __builtin_inff();
    double pm[M::MAXS];
    for (int s = 0; s < M::MAXS; ++s) pm[s] = 0.0;
    float* Xc = X[cur];
    float* Xn = X[cur ^ 1];
    // The model's particle-independent sub-step values (thread k: sub-step k,
    // its start time summed in the order the step loop adds) and observation
    // term (one thread of another wave) are computed once here: all of them
    // with one parameter set for all particles (mode 1), their time- and
    // data-only parts with per-particle parameters (mode 2, IF2).  The previous
    // observation's reductions end with a barrier, so the tables are free to
    // overwrite.
    constexpr int kHoist = !M::HOIST ? 0 : (ALGO == ALGO_PER ? 2 : 1);
    using HT = HoistTables<M, kHoist>;
    __shared__ typename HT::Sub s_sub[kHoist ? kMaxSub : 1];
    __shared__ typename HT::Obs s_obs;
    const bool hoisted = kHoist && nst <= kMaxSub;
    if constexpr (kHoist != 0) {
      if (hoisted) {
        if (tid < nst) {
          float tk = tstart;
          for (int k = 0; k < tid; ++k) tk = tk + dts[ti0 + k];
          if constexpr (kHoist == 1) s_sub[tid] = M::substep(P, tk, dts[ti0 + tid]);
          else s_sub[tid] = M::tab_sub(tk);
        }
        if (tid == WAVE) {
          if constexpr (kHoist == 1) s_obs = M::obs(P, y);
          else s_obs = M::tab_obs(y);
        }
        __syncthreads();
      }
    }

    for (int j = tid; j < J; j += B) {
      int src = have_anc ? ancestor_seg(Cc, tab, J, ures, total, j) : j;
      float x[WM];
      MP_UNROLL for (int s = 0; s < WM; ++s) if (s < W) x[s] = Xc[(int64_t)s * J + src];
      const float wprev = uniform_w ? wconst : lw[src];
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s)
        if ((md.accum_mask >> s) & 1u) x[s] = 0.f;
      typename M::P Pj;
      if (ALGO == ALGO_PER) {
        float th[M::MAXP];
        MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < d) {
          float z = normal_at(kpert, (uint64_t)j * d + p);
          x[S + p] = x[S + p] + (cft * a.sigmas[p]) * z;
          th[p] = x[S + p];
        }
        if (!md.identity_trans) M::from_est(md, th);
        M::prepare(md, th, Pj);
      } else {
        Pj = P;
      }
      float tt = tstart;
      float meas;
      if constexpr (kHoist != 0) {
        if (hoisted) {
          for (int k = 0; k < nst; ++k) {
            const Key sk = ld_key(sch, base + k);
            const float dtk = dts[ti0 + k];
            if constexpr (kHoist == 1) M::step_sub(Pj, s_sub[k], split_at(sk, (uint64_t)j), dtk, x);
            else M::step_tab(Pj, s_sub[k], split_at(sk, (uint64_t)j), dtk, x);
          }
          if constexpr (kHoist == 1) meas = M::dmeas_obs(Pj, s_obs, y, x);
          else meas = M::dmeas_obs(Pj, M::obs_tab(Pj, s_obs, y), y, x);
        }
      }
      if (!hoisted) {
        for (int k = 0; k < nst; ++k) {
          const Key sk = ld_key(sch, base + k);
          const float dt = dts[ti0 + k];
          m_step<M>(md, Pj, sk, (uint64_t)j, tt, dt, x, cov_row(a.ser, u, ti0 + k));
          tt = tt + dt;
        }
        meas = m_dmeas<M>(md, Pj, y, x, ty, cov_row(a.ser, u, ti0 + nst));
      }
      if (ALGO == ALGO_PER) {
        if (meas != meas) meas = nan_sub;
        else if (meas == __builtin_inff()) meas = 3.40282347e+38f;
        else if (meas == -__builtin_inff()) meas = -3.40282347e+38f;
      }
      // MOP keeps the measurement density alone: the resampling weights are
      // normalize(meas) (mop.py:120) and wP is common to all particles.
      const float w = (ALGO == ALGO_MOP) ? meas : (wprev + meas);
      MP_UNROLL for (int s = 0; s < WM; ++s) if (s < W) Xn[(int64_t)s * J + j] = x[s];
      lw[j] = w;
      lmax = nanmax(lmax, w);
      lmin = nanmin(lmin, w);
      if (a.flags & F_PMEAN)
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) pm[s] += (double)x[s];
    }
    block_maxmin(lmax, lmin, red);
    const bool collapsed = (lmax == -__builtin_inff());
    // fixed-point weights and their prefix sums: the resampler's input and the
    // exact total of the log-likelihood term
    total = block_fixed_scan_seg(lw, Cc, J, lmax, tab);
    double llt = step_loglik(lmax, total, shift_for(J));
    // MOP: LSE(wP + m) - LSE(wP) with wP common = max(m) + log sum exp(m - max) - log J
    if (ALGO == ALGO_MOP) llt = llt - (double)logJ;
    loglik += llt;
    const float llf = (float)llt;
    if (tid == 0 && (a.flags & F_CLL)) a.cll[(int64_t)f * a.ser.Tmax + t] = llf;

    bool resample;
    if (ALGO == ALGO_MOP) {
      resample = true;
    } else {
      float nmax = collapsed ? -logJ : lmax - llf;
      float nmin = collapsed ? -logJ : lmin - llf;
      resample = (nmax - nmin) > a.logthresh;
    }
    if (a.flags & (F_ESS | F_FMEAN | F_PMEAN)) {
      double acc[1 + M::MAXS + M::MAXS];
      for (int k = 0; k < 1 + 2 * M::MAXS; ++k) acc[k] = 0.0;
      if (a.flags & (F_ESS | F_FMEAN)) {
        for (int j = tid; j < J; j += B) {
          float nw = collapsed ? -logJ : lw[j] - llf;
          float e = expf(nw);
          acc[0] += (double)(e * e);
          if (a.flags & F_FMEAN)
            MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) acc[1 + s] += (double)(Xn[(int64_t)s * J + j] * e);
        }
      }
      for (int s = 0; s < M::MAXS; ++s) acc[1 + M::MAXS + s] = pm[s];
      block_sum_n<1 + 2 * M::MAXS>(acc, red);
      if (tid == 0) {
        const int64_t o = (int64_t)f * a.ser.Tmax + t;
        if (a.flags & F_ESS) a.ess[o] = (float)(1.0 / acc[0]);
        MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) {
          if (a.flags & F_FMEAN) a.fmean[o * S + s] = (float)acc[1 + s];
          if (a.flags & F_PMEAN) a.pmean[o * S + s] = (float)(acc[1 + M::MAXS + s] / (double)J);
        }
      }
    }

    const bool last = (t == T - 1);
    if (resample && !(last && ALGO != ALGO_PER)) {
      // resampling weights: PF/PER use the normalised weights, MOP the
      // measurement densities alone (mop.py:120); both equal lw up to a
      // common shift, which the fixed-point weights (C, total) factor out.
      const Key rk = ld_key(sch, base + nst);
      ures = uniform01(bits_at(rk, 0));
      have_anc = true;
      uniform_w = true;
      wconst = 0.0f - logJ;
      if (ALGO == ALGO_MOP) mop_c = mop_wP;
    } else {
      have_anc = false;
      uniform_w = false;
      for (int j = tid; j < J; j += B) lw[j] = collapsed ? -logJ : lw[j] - llf;
      if (ALGO == ALGO_MOP) mop_c = mop_wP;
      __syncthreads();
    }
    cur ^= 1;
  }

  if (ALGO == ALGO_PER) {
    double mean[M::MAXP];
    for (int p = 0; p < M::MAXP; ++p) mean[p] = 0.0;
    for (int j = tid; j < J; j += B) {
      int src = have_anc ? ancestor_seg(Cc, tab, J, ures, total, j) : j;
      MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < d) {
        float v = X[cur][(int64_t)(S + p) * J + src];
        a.theta_out[((int64_t)f * J + j) * d + p] = v;
        mean[p] += (double)v;
      }
    }
    if (track)
      for (int j = tid; j < J; j += B) {
        int src = have_anc ? ancestor_seg(Cc, tab, J, ures, total, j) : j;
        a.ancestry[(int64_t)f * J + j] = __float_as_int(X[cur][(int64_t)(S + d) * J + src]);
      }
    block_sum_n<M::MAXP>(mean, red);
    if (tid == 0)
      MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < d) a.theta_mean[(int64_t)f * d + p] = (float)(mean[p] / (double)J);
  }
  if (tid == 0) a.loglik[f] = loglik;
  __syncthreads();
}

// Persistent workgroups take filters from the queue, longest unit first
// (order[] from k_order), so the uneven units of a panel balance over the CUs.
// A filter's result does not depend on which workgroup runs it.  Global
// scratch is indexed by the filter's queue position, which every workgroup
// takes once.
// LDSR: the filter's buffers live in LDS.  A compile-time choice, so their
// accesses compile to LDS (or global) instructions rather than generic flat ones.
template <class M, int ALGO, int NT, bool LDSR>
__global__ __launch_bounds__(NT) void k_batched(const BatchArgs a) {
  extern __shared__ __align__(256) char smem[];
  __shared__ int next;
  double* red = reinterpret_cast<double*>(smem);
  uint64_t* tab = reinterpret_cast<uint64_t*>(smem + kRedBytes);
  char* ldsbuf = smem + kRedBytes + batched_tab_bytes(a.J, blockDim.x);
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(a.queue, 1);
    __syncthreads();
    const int q = next;
    __syncthreads();
    if (q >= a.R) break;
    const int f = a.order[q];
    if constexpr (LDSR) run_filter<M, ALGO>(a, f, ldsbuf, red, tab);
    else run_filter<M, ALGO>(a, f, a.ws + (int64_t)q * a.ws_stride, red, tab);
  }
}

// Filter order for the work queue: filters sorted by their unit's length,
// longest first, ties by index (a rank per filter; R is at most a few
// thousand).  Also resets the queue counter, so a captured launch sequence
// needs no memset.
__global__ void k_order(SeriesSet ser, const int32_t* __restrict__ unit, int R, int32_t* __restrict__ order,
                        int32_t* __restrict__ queue) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f == 0) *queue = 0;
  if (f >= R) return;
  auto len = [&](int g) {
    const int u = unit ? unit[g] : 0;
    return ser.obs_off[u + 1] - ser.obs_off[u];
  };
  const int64_t Tf = len(f);
  int rank = 0;
  for (int g = 0; g < R; ++g) {
    const int64_t Tg = len(g);
    rank += (Tg > Tf) || (Tg == Tf && g < f);
  }
  order[rank] = f;
}


// Host launchers ------------------------------------------------------------

// Workgroup sizes of the batched form: 256 when several filters share a CU,
// up to 768 (12 waves) when one large filter per CU must fill it alone.
template <class M, int ALGO>
static hipError_t launch_batched_t(const BatchArgs& a, int block, int grid, size_t lds, hipStream_t st) {
  switch (block) {
    case 256:
      if (a.lds_resident) mp_launch(k_batched<M, ALGO, 256, true>, dim3(grid), dim3(256), lds, st, a);
      else mp_launch(k_batched<M, ALGO, 256, false>, dim3(grid), dim3(256), lds, st, a);
      break;
    case 512:
      if (a.lds_resident) mp_launch(k_batched<M, ALGO, 512, true>, dim3(grid), dim3(512), lds, st, a);
      else mp_launch(k_batched<M, ALGO, 512, false>, dim3(grid), dim3(512), lds, st, a);
      break;
    case 768:
      if (a.lds_resident) mp_launch(k_batched<M, ALGO, 768, true>, dim3(grid), dim3(768), lds, st, a);
      else mp_launch(k_batched<M, ALGO, 768, false>, dim3(grid), dim3(768), lds, st, a);
      break;
    case 1024:
      // 16 waves at 4 per SIMD (128 VGPRs): the models whose particles share
      // the parameters through LDS tables (batched_block_fits_1024)
      if constexpr (M::HOIST && ALGO != ALGO_PER) {
        if (a.lds_resident) mp_launch(k_batched<M, ALGO, 1024, true>, dim3(grid), dim3(1024), lds, st, a);
        else mp_launch(k_batched<M, ALGO, 1024, false>, dim3(grid), dim3(1024), lds, st, a);
        break;
      } else {
        return hipErrorInvalidValue;
      }
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int ALGO>
static hipError_t launch_batched_algo(const BatchArgs& a, int block, int grid, size_t lds, hipStream_t st) {
  switch (a.md.kind) {
    case MODEL_LG:
      if (a.md.dx == 1 && a.md.dy == 1) return launch_batched_t<LG1, ALGO>(a, block, grid, lds, st);
      if (a.md.dx == 2 && a.md.dy == 2) return launch_batched_t<LG2, ALGO>(a, block, grid, lds, st);
      if (a.md.dx <= 4 && a.md.dy <= 4) return launch_batched_t<LG4, ALGO>(a, block, grid, lds, st);
      // dimensions 5..8: 256-thread workgroups only (the plan fixes them)
      if (block != 256) return hipErrorInvalidValue;
      if (a.lds_resident) mp_launch(k_batched<LG8, ALGO, 256, true>, dim3(grid), dim3(256), lds, st, a);
      else mp_launch(k_batched<LG8, ALGO, 256, false>, dim3(grid), dim3(256), lds, st, a);
      return hipGetLastError();
    case MODEL_SV: return launch_batched_t<SV, ALGO>(a, block, grid, lds, st);
    case MODEL_SEIR: return launch_batched_t<SEIR, ALGO>(a, block, grid, lds, st);
    case MODEL_SPX: return launch_batched_t<SPX, ALGO>(a, block, grid, lds, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batched(int algo, const BatchArgs& a, int block, int grid, size_t lds, hipStream_t st) {
  switch (algo) {
    case ALGO_PF: return launch_batched_algo<ALGO_PF>(a, block, grid, lds, st);
    case ALGO_MOP: return launch_batched_algo<ALGO_MOP>(a, block, grid, lds, st);
    case ALGO_PER: return launch_batched_algo<ALGO_PER>(a, block, grid, lds, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_order(const SeriesSet& ser, const int32_t* unit, int R, int32_t* order, int32_t* queue,
                        hipStream_t st) {
  mp_launch(k_order, dim3((R + 255) / 256), dim3(256), 0, st, ser, unit, R, order, queue);
  return hipGetLastError();
}

hipError_t launch_sched(int algo, const uint32_t* keys, const SeriesSet& ser, const int32_t* unit,
                        int64_t sched_len, uint32_t* sched, int R, hipStream_t st) {
  if (algo == ALGO_PER) {
    dim3 grid((ser.Tmax + 1 + 127) / 128, R);
    mp_launch(k_sched_per, grid, dim3(128), 0, st, keys, ser, unit, sched_len, sched, R);
  } else {
    mp_launch(k_sched_pf, dim3((4 * R + 63) / 64), dim3(64), 0, st, keys, ser, unit, sched_len, sched, R);
  }
  return hipGetLastError();
}

size_t batched_small_lds(int J, int block) { return kRedBytes + batched_tab_bytes(J, block); }

}  // namespace mp
