// Plain form: one filter at a time, its particles spread over the whole GPU,
// two launches per observation (the shape of the reference's per-step XLA
// loop body, pfilter.py:97-182):
//
//  k_plain_prop  -- ancestor search (systematic comb against the previous
//                   step's prefix sums, staged per tile in LDS) fused with the
//                   gather, the Euler sub-steps, the measurement density and the
//                   log-weight; per-tile max/min.
//  k_plain_scan  -- fixed-point weights and a one-pass inclusive prefix sum
//                   across tiles with decoupled look-back (bounded waits: a
//                   tile computes a late predecessor's aggregate itself); the
//                   tile with the last index finishes the step from the exact
//                   total: the log-likelihood term (FP64 sum over t), the
//                   resampling decision and the comb's uniform.
//
// Cross-workgroup traffic uses only relaxed agent-scope atomics on single
// 64-bit words (look-back state and prefix packed together): on a multi-XCD
// GPU an agent-scope release/acquire fence writes back / invalidates the whole
// L2, which is what this design avoids.  Everything else a kernel reads was
// written by an earlier launch.
#include "blockops.hpp"
#include "engine.hpp"
#include "mech.hpp"

namespace mp {

constexpr uint64_t kStateShift = 62;
constexpr uint64_t kAggState = 1ull << kStateShift, kInclState = 2ull << kStateShift;
constexpr uint64_t kValueMask = (1ull << kStateShift) - 1;
constexpr int kWin = 2048;  // LDS search window (uint64 entries)
constexpr int kMaxScanRows = 16;  // rows of 256 particles per scan tile, at most
constexpr int kLookback = 4;      // look-back loads per lane and probe

__device__ __forceinline__ uint64_t ld_relaxed(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Chunk of propagate block b when a scan tile holds q chunks: blocks go to the
// kXcds XCDs round-robin, so within each group of kXcds * q blocks, block r
// (on XCD r % kXcds) takes chunk q * (r % kXcds) + r / kXcds; the q chunks of
// scan tile k of the group then all run on XCD k % kXcds, the scan tile's own.
__device__ __forceinline__ int xcd_chunk(int b, int q) {
  const int G = kXcds * q, g = b / G, r = b - g * G;
  return g * G + q * (r % kXcds) + r / kXcds;
}

// The step scalars of observation t (PlainState's step fields).
__device__ void load_step(const PlainArgs& a, int t, PlainState& s) {
  const int64_t obs = a.ser.obs_off[a.unit] + t;
  const int ti0 = a.ser.tidx0[obs], nst = a.ser.nstep[obs];
  const int64_t base = ti0 + (int64_t)t;
  s.obs = obs;
  s.base = base;
  s.ti0 = ti0;
  s.nst = nst;
  s.tstart = t == 0 ? a.ser.t0 : a.ser.times[obs - 1];
  s.ty = a.ser.times[obs];
  s.dt0 = nst > 0 ? a.ser.dt_ext[a.ser.dt_off[a.unit] + ti0] : 0.0f;
  s.k0 = a.sched[2 * base];
  s.k1 = a.sched[2 * base + 1];
  s.r0 = a.sched[2 * (base + nst)];
  s.r1 = a.sched[2 * (base + nst) + 1];
}

template <class M>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void k_plain_prop(const PlainArgs a) {
  __shared__ double red[32];
  __shared__ int64_t s_win[2];
  __shared__ uint64_t s_C[kWin];
  const ModelDesc md = a.md;
  const int J = a.J, S = M::FIXED_S ? M::MAXS : md.S, t = a.t, tid = threadIdx.x, B = blockDim.x;
  // the step state, with this step's scalars (row, grid, keys) in it
  const PlainState st = *a.st;
  const int ti0 = st.ti0;
  const int nst = st.nst;
  const float* dts = a.ser.dt_ext + a.ser.dt_off[a.unit];
  const float ty = st.ty;
  const float tstart = st.tstart;
  const float* y = a.ser.ys + st.obs * md.dy;
  const int64_t base = st.base;
  const float fJ = (float)J;
  const float logJ = d_log(fJ);

  // parameters prepared once per filter (k_plain_prepare), staged in LDS
  __shared__ typename M::P s_P;
  if (tid == 0) s_P = *reinterpret_cast<const typename M::P*>(a.params);
  __syncthreads();
  const typename M::P& P = s_P;

  const float* Xc = (t & 1) ? a.X1 : a.X0;
  float* Xn = (t & 1) ? a.X0 : a.X1;
  const float wconst = t == 0 ? d_log(1.0f / fJ) : st.wconst;
  const bool uniform_w = t == 0 ? true : st.uniform_w;
  const bool have_anc = t == 0 ? false : st.have_anc;

  float lmax = -__builtin_inff(), lmin = __builtin_inff();
  const int chunk = a.xcd_q > 0 ? xcd_chunk(blockIdx.x, a.xcd_q) : (int)blockIdx.x;
  const int64_t tile0 = (int64_t)chunk * a.prop_len;
  const int64_t tile1 = min((int64_t)J, tile0 + (int64_t)a.prop_len);
  if (tile0 >= J) {  // a trailing block without particles
    if (tid == 0) {
      a.pmax[blockIdx.x] = lmax;
      a.pmin[blockIdx.x] = lmin;
    }
    return;
  }
  // Ancestor search window of this tile.  A(t) = #{i : C_i <= t} is
  // non-decreasing in the comb position, so every j of the tile has
  // A(t_j) in [A(t_tile0), A(t_tile1-1)]; that slice of C is staged in LDS.
  // Waves 0 and 1 find the two ends of the window concurrently.
  const int wv = tid >> 6;
  if (have_anc && wv < 2) {
    const int64_t jj = wv == 0 ? tile0 : tile1 - 1;
    const int64_t v = wave_upper_bound_u64(a.C, 0, J, comb_threshold(st.ures, jj, J, st.total));
    if ((tid & 63) == 0) s_win[wv] = v;
  }
  // Models whose step noise depends on the particle's key alone draw it for
  // the first kPreRows rows here: the other waves do this VALU work while
  // waves 0 and 1 wait on the search's dependent loads.
  constexpr int kPreRows = 2;
  constexpr int NZ = M::NOISE > 0 ? M::NOISE : 1;
  const bool pre = M::NOISE > 0 && nst == 1;
  float zpre[kPreRows][NZ];
  if (pre) {
    const Key sk = uniform_key(st.k0, st.k1);
    MP_UNROLL for (int r = 0; r < kPreRows; ++r) {
      const int64_t j = tile0 + (int64_t)r * B + tid;
      if (r < a.rows_prop && j < tile1) m_draw<M>(md, P, sk, (uint64_t)j, zpre[r]);
    }
  }
  if (have_anc) __syncthreads();
  const int64_t wlo = have_anc ? s_win[0] : 0, whi = have_anc ? s_win[1] : 0;
  const bool in_lds = have_anc && (whi - wlo) <= kWin;
  if (in_lds) {
    for (int64_t i = tid; i < whi - wlo; i += B) s_C[i] = a.C[wlo + i];
    __syncthreads();
  }
  auto row = [&](int r, bool use_z, const float (&zp)[NZ]) {
    const int64_t j = tile0 + (int64_t)r * B + tid;
    if (j >= tile1) return;
    float x[M::MAXS];
    int64_t src = j;
    if (t == 0) {
      m_rinit<M>(md, P, split_at(Key{a.fkey[0], a.fkey[1]}, 1 + (uint64_t)j), a.ser.t0, x,
                 cov_row(a.ser, a.unit, 0));
    } else {
      if (have_anc) {
        const uint64_t tj = comb_threshold(st.ures, j, J, st.total);
        const int64_t an = in_lds ? wlo + (whi > wlo ? upper_bound_fixed(s_C, (int)(whi - wlo), tj) : 0)
                                  : upper_bound_u64(a.C, wlo, whi, tj);
        src = an < J ? an : J - 1;
      }
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = Xc[(int64_t)s * J + src];
    }
    float wprev = uniform_w ? wconst : (st.collapsed ? -logJ : a.lw[j] - st.llf);
    MP_UNROLL for (int s = 0; s < M::MAXS; ++s)
      if ((md.accum_mask >> s) & 1u) x[s] = 0.f;
    float tt = tstart;
    if (use_z) {
      m_step_z<M>(md, P, zp, tt, st.dt0, x, cov_row(a.ser, a.unit, ti0));
    } else {
      for (int k = 0; k < nst; ++k) {
        const Key sk = uniform_key(a.sched[2 * (base + k)], a.sched[2 * (base + k) + 1]);
        const float dt = dts[ti0 + k];
        m_step<M>(md, P, sk, (uint64_t)j, tt, dt, x, cov_row(a.ser, a.unit, ti0 + k));
        tt = tt + dt;
      }
    }
    const float w = wprev + m_dmeas<M>(md, P, y, x, ty, cov_row(a.ser, a.unit, ti0 + nst));
    MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) Xn[(int64_t)s * J + j] = x[s];
    a.lw[j] = w;
    lmax = nanmax(lmax, w);
    lmin = nanmin(lmin, w);
  };
  MP_UNROLL for (int r = 0; r < kPreRows; ++r)
    if (r < a.rows_prop) row(r, pre, zpre[r]);
  for (int r = kPreRows; r < a.rows_prop; ++r) row(r, false, zpre[0]);
  block_maxmin(lmax, lmin, red);
  if (tid == 0) {
    a.pmax[blockIdx.x] = lmax;
    a.pmin[blockIdx.x] = lmin;
  }
}

// Aggregate of scan tile `idx` (the sum of its particles' fixed-point weights),
// computed by one wave: what that tile publishes, computed here instead when it
// has not published in time (a tile whose workgroup is not running yet).
__device__ uint64_t wave_tile_aggregate(const float* lw, int64_t idx, int64_t tile_len, int64_t J, float m,
                                        int sh) {
  const int lane = threadIdx.x & 63;
  const int64_t j0 = idx * tile_len, j1 = min(J, j0 + tile_len);
  uint64_t v = 0;
  for (int64_t j = j0 + lane; j < j1; j += 64) v += particle_fixed_weight(lw[j], m, sh);
  return wave_sum_u64(v);
}

template <class M>
__global__ __launch_bounds__(1024) void k_plain_scan(const PlainArgs a) {
  __shared__ double red[32];
  __shared__ uint64_t tab[65];  // wave totals (<= 16) and the tile total
  __shared__ uint64_t s_prefix;
  const int J = a.J, tid = threadIdx.x, B = blockDim.x, lane = tid & 63, w = tid >> 6;
  const int nw = (B + 63) >> 6;
  const int ntiles = a.ntiles;
  const int par = a.t & 1;
  uint64_t* lb = a.lb + (int64_t)par * ntiles;
  const int sh = shift_for(J);
  const int tile = blockIdx.x;
  const int rows = a.rows;
  const int64_t tile0 = (int64_t)tile * rows * B;
  // Each thread owns `rows` consecutive particles j0 .. j0 + rows - 1: it
  // loads them as float4 vectors, scans them serially and takes one wave scan
  // of its total (rather than one wave scan per row).
  const int64_t j0 = tile0 + (int64_t)tid * rows;
  const bool full = j0 + rows <= J;  // all of the thread's particles exist
  const bool vec = full && (rows & 3) == 0;

  // the thread's log-weights, loads issued first so that they are in flight
  // together with the per-block maxima below
  float lwv[kMaxScanRows];
  if (vec) {
    const float4* p4 = reinterpret_cast<const float4*>(a.lw + j0);
#pragma unroll
    for (int q = 0; q < kMaxScanRows / 4; ++q) {
      if (4 * q < rows) {
        const float4 v = p4[q];
        lwv[4 * q] = v.x;
        lwv[4 * q + 1] = v.y;
        lwv[4 * q + 2] = v.z;
        lwv[4 * q + 3] = v.w;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < kMaxScanRows; ++r) lwv[r] = (r < rows && j0 + r < J) ? a.lw[j0 + r] : -__builtin_inff();
  }
  // The finishing tile's scalars, loaded now so that their dependent loads
  // overlap the scan: the comb uniform from this step's resampling key, and
  // the next observation's row, grid and keys.
  const bool finisher = tile == ntiles - 1 && tid == 0;
  float ures_next = 0.0f;
  double loglik_prev = 0.0;
  PlainState nx{};
  if (finisher) {
    const PlainState cur = *a.st;
    loglik_prev = cur.loglik;
    ures_next = uniform01(bits_at(Key{cur.r0, cur.r1}, 0));
    if (a.t + 1 < a.T) load_step(a, a.t + 1, nx);
  }
  // global max / min of the log-weights from the per-block values, eight
  // independent loads in flight per thread
  float m = -__builtin_inff(), mn = __builtin_inff();
  for (int i0 = 0; i0 < a.nprop; i0 += 8 * B) {
    float vx[8], vn[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + k * B + tid;
      vx[k] = i < a.nprop ? a.pmax[i] : -__builtin_inff();
      vn[k] = i < a.nprop ? a.pmin[i] : __builtin_inff();
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      m = nanmax(m, vx[k]);
      mn = nanmin(mn, vn[k]);
    }
  }
  block_maxmin(m, mn, red);

  // fixed-point weights, their running sum over the thread's particles, and
  // one wave scan of the thread totals
  uint64_t loc[kMaxScanRows];
  uint64_t run = 0;
#pragma unroll
  for (int r = 0; r < kMaxScanRows; ++r) {
    if (r < rows) {
      run += (j0 + r < J) ? particle_fixed_weight(lwv[r], m, sh) : 0;
      loc[r] = run;
    }
  }
  const uint64_t winc = wave_incl_scan_u64(run);
  const uint64_t texcl = winc - run;
  if (lane == 63) tab[w] = winc;
  __syncthreads();
  // wave 0: exclusive scan of the wave totals, then publish the tile's
  // aggregate
  const int n = nw;
  if (w == 0) {
    const uint64_t v = lane < n ? tab[lane] : 0;
    const uint64_t inc = wave_incl_scan_u64(v);
    if (lane < n) tab[lane] = inc - v;
    const uint64_t aggv = __shfl(inc, WAVE - 1, WAVE);
    if (lane == 0) {
      tab[n] = aggv;
      st_relaxed(&lb[tile], (tile == 0 ? kInclState : kAggState) | aggv);
    }
    // Decoupled look-back, kLookback * 64 predecessors per probe with all of
    // their loads in flight together: lane l inspects tiles base - l - 64 k.
    // The nearest predecessor with an inclusive prefix ends the walk; the
    // aggregates in between are summed with a wave reduction.  Up to
    // kLookback * 64 tiles this is one round trip once the predecessors have
    // published.  Tiles are block indices, so a predecessor may not be running
    // yet (its workgroup not dispatched while this one waits): after
    // a.scan_spin probes without progress the wave computes the missing
    // aggregates itself from the log-weights.  Every wait is therefore bounded
    // and no dispatch order is assumed.
    uint64_t prefix = 0;
    if (tile > 0) {
      int bse = tile - 1;
      int spins = 0;
      while (true) {
        uint64_t word[kLookback];
#pragma unroll
        for (int k = 0; k < kLookback; ++k) {
          const int idx = bse - lane - 64 * k;
          word[k] = idx >= 0 ? ld_relaxed(&lb[idx]) : kInclState;
        }
        // offset (lane + 64 k) of the nearest inclusive predecessor
        int first = 64 * kLookback;
#pragma unroll
        for (int k = kLookback - 1; k >= 0; --k) {
          const uint64_t im = __ballot((word[k] >> kStateShift) == 2);
          if (im) first = 64 * k + __ffsll((unsigned long long)im) - 1;
        }
        uint64_t nr[kLookback];
        bool ready = true;
#pragma unroll
        for (int k = 0; k < kLookback; ++k) {
          nr[k] = __ballot((word[k] >> kStateShift) == 0 && 64 * k + lane <= first);
          ready = ready && nr[k] == 0;
        }
        if (!ready) {
          if (spins < a.scan_spin) {
            ++spins;
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
#pragma unroll
          for (int k = 0; k < kLookback; ++k) {
            uint64_t msk = nr[k];
            while (msk) {
              const int l = __ffsll((unsigned long long)msk) - 1;
              msk &= msk - 1;
              const uint64_t agg = wave_tile_aggregate(a.lw, bse - l - 64 * k, (int64_t)rows * B, J, m, sh);
              if (lane == l) word[k] = kAggState | agg;
            }
          }
        }
        spins = 0;
        uint64_t v = 0;
#pragma unroll
        for (int k = 0; k < kLookback; ++k) v += 64 * k + lane <= first ? (word[k] & kValueMask) : 0;
        prefix += wave_sum_u64(v);
        if (first < 64 * kLookback) break;
        bse -= 64 * kLookback;
      }
      if (lane == 0) st_relaxed(&lb[tile], kInclState | (prefix + aggv));
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  const uint64_t base = prefix + tab[w] + texcl;
  if (full && (rows & 1) == 0) {
    ulonglong2* p2 = reinterpret_cast<ulonglong2*>(a.C + j0);
#pragma unroll
    for (int q = 0; q < kMaxScanRows / 2; ++q)
      if (2 * q < rows) p2[q] = ulonglong2{base + loc[2 * q], base + loc[2 * q + 1]};
  } else {
#pragma unroll
    for (int r = 0; r < kMaxScanRows; ++r)
      if (r < rows && j0 + r < J) a.C[j0 + r] = base + loc[r];
  }
  // The tile holding the last index finishes the step: its inclusive prefix is
  // the exact total of the fixed-point weights.  It touches only the step
  // state and the other parity's words, which no tile of this launch reads.
  if (tile != ntiles - 1) return;
  if (tid == 0) {
    const uint64_t total = prefix + tab[n];
    PlainState* st = a.st;
    const float fJ = (float)J, logJ = d_log(fJ);
    const bool collapsed = (m == -__builtin_inff());
    const double llt = step_loglik(m, total, sh);
    const float llf = (float)llt;
    st->loglik = (a.t == 0 ? 0.0 : loglik_prev) + llt;
    if (a.cll) a.cll[a.t] = llf;
    const float nmax = collapsed ? -logJ : m - llf;
    const float nmin = collapsed ? -logJ : mn - llf;
    const bool res = (nmax - nmin) > a.logthresh;
    st->llf = llf;
    st->collapsed = collapsed;
    st->have_anc = res;
    st->uniform_w = res;
    st->wconst = 0.0f - logJ;
    st->total = total;
    st->ures = ures_next;
    if (a.t + 1 < a.T) {
      st->obs = nx.obs;
      st->base = nx.base;
      st->ti0 = nx.ti0;
      st->nst = nx.nst;
      st->tstart = nx.tstart;
      st->ty = nx.ty;
      st->dt0 = nx.dt0;
      st->k0 = nx.k0;
      st->k1 = nx.k1;
      st->r0 = nx.r0;
      st->r1 = nx.r1;
    }
    if (a.t == a.T - 1) *a.loglik_out = (a.t == 0 ? 0.0 : loglik_prev) + llt;
  }
  // clear the next step's look-back words
  uint64_t* lbn = a.lb + (int64_t)(par ^ 1) * ntiles;
  for (int i = tid; i < ntiles; i += B) lbn[i] = 0;
}

// Diagnostics of step t (pfilter.py:146-156), after its scan: ESS = 1 / sum
// exp(2 nw), filter mean = sum x exp(nw), prediction mean = mean x, with nw =
// lw - llf (-log J after a collapse) the normalised weights before
// resampling; float32 terms as the batched form forms them, summed in float64.
// Each block writes its partial sums; k_plain_diag_final adds them in block
// order (a fixed order: the same bits for every run).
template <class M>
__global__ __launch_bounds__(256) void k_plain_diag(const PlainArgs a) {
  __shared__ double red[kRedUsable / 8 + 1];
  constexpr int NV = 1 + 2 * M::MAXS;
  static_assert(NV <= kDiagWords, "diagnostic words");
  const int J = a.J, S = M::FIXED_S ? M::MAXS : a.md.S, tid = threadIdx.x, B = blockDim.x;
  const float* Xn = (a.t & 1) ? a.X0 : a.X1;
  const PlainState st = *a.st;
  const float logJ = d_log((float)J);
  double acc[NV];
  MP_UNROLL for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * B + tid; j < J; j += (int64_t)gridDim.x * B) {
    const float nw = st.collapsed ? -logJ : a.lw[j] - st.llf;
    const float e = expf(nw);
    acc[0] += (double)(e * e);
    MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) {
      const float x = Xn[(int64_t)s * J + j];
      acc[1 + s] += (double)(x * e);
      acc[1 + M::MAXS + s] += (double)x;
    }
  }
  block_sum_n<NV>(acc, red);
  if (tid == 0) MP_UNROLL for (int k = 0; k < NV; ++k) a.diag_part[(int64_t)blockIdx.x * kDiagWords + k] = acc[k];
}

template <class M>
__global__ __launch_bounds__(64) void k_plain_diag_final(const PlainArgs a) {
  constexpr int NV = 1 + 2 * M::MAXS;
  const int k = threadIdx.x, S = M::FIXED_S ? M::MAXS : a.md.S;
  if (k >= NV) return;
  double v = 0.0;
  for (int b = 0; b < a.ndiag; ++b) v += a.diag_part[(int64_t)b * kDiagWords + k];
  const int t = a.t;
  if (k == 0) {
    if (a.ess) a.ess[t] = (float)(1.0 / v);
  } else if (k <= M::MAXS) {
    const int s = k - 1;
    if (a.fmean && s < S) a.fmean[(int64_t)t * S + s] = (float)v;
  } else {
    const int s = k - 1 - M::MAXS;
    if (a.pmean && s < S) a.pmean[(int64_t)t * S + s] = (float)(v / (double)a.J);
  }
}

// Prepared model parameters of each filter (one thread per filter).
template <class M>
__global__ void k_plain_prepare(ModelDesc md, const float* theta, int64_t stride, int R, int flags, char* out,
                                int64_t pstride) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= R) return;
  float th[M::MAXP];
  MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < md.d) th[p] = theta[f * stride + p];
  if ((flags & F_SHOULD_TRANS) && !md.identity_trans) M::from_est(md, th);
  typename M::P P;
  M::prepare(md, th, P);
  *reinterpret_cast<typename M::P*>(out + f * pstride) = P;
}

template <class M>
static hipError_t plain_prepare_t(const ModelDesc& md, const float* theta, int64_t stride, int R, int flags,
                                  char* out, int64_t pstride, hipStream_t s) {
  static_assert(sizeof(typename M::P) <= 512, "prepared parameters exceed their slot");
  mp_launch(k_plain_prepare<M>, dim3((R + 63) / 64), dim3(64), 0, s, md, theta, stride, R, flags, out,
                     pstride);
  return hipGetLastError();
}

hipError_t plain_prepare(const ModelDesc& md, const float* theta, int64_t stride, int R, int flags, char* out,
                         int64_t pstride, hipStream_t s) {
  switch (md.kind) {
    case MODEL_LG:
      if (md.dx == 1 && md.dy == 1) return plain_prepare_t<LG1>(md, theta, stride, R, flags, out, pstride, s);
      if (md.dx == 2 && md.dy == 2) return plain_prepare_t<LG2>(md, theta, stride, R, flags, out, pstride, s);
      return plain_prepare_t<LG4>(md, theta, stride, R, flags, out, pstride, s);
    case MODEL_SV: return plain_prepare_t<SV>(md, theta, stride, R, flags, out, pstride, s);
    case MODEL_SEIR: return plain_prepare_t<SEIR>(md, theta, stride, R, flags, out, pstride, s);
    case MODEL_SPX: return plain_prepare_t<SPX>(md, theta, stride, R, flags, out, pstride, s);
  }
  return hipErrorInvalidValue;
}

// Start of a filter: clear the look-back words and the step state, with the
// first observation's scalars in it.
__global__ void k_plain_reset(const PlainArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.ntiles) {
    a.lb[i] = 0;
    a.lb[a.ntiles + i] = 0;
  }
  if (i == 0) {
    PlainState z{};
    if (a.T > 0) load_step(a, 0, z);
    *a.st = z;
  }
}

hipError_t plain_reset(const PlainArgs& a, hipStream_t s) {
  mp_launch(k_plain_reset, dim3((a.ntiles + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

// which = 0: propagate kernel, 1: scan kernel, 2: step diagnostics
template <class M>
static hipError_t plain_kernel_t(int which, const PlainArgs& a, int ntiles, int threads, hipStream_t s) {
  if (which == 0) {
    mp_launch(k_plain_prop<M>, dim3(a.nprop), dim3(threads), 0, s, a);
  } else if (which == 1) {
    mp_launch(k_plain_scan<M>, dim3(ntiles), dim3(threads), 0, s, a);
  } else {
    mp_launch(k_plain_diag<M>, dim3(a.ndiag), dim3(256), 0, s, a);
    mp_launch(k_plain_diag_final<M>, dim3(1), dim3(64), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t plain_kernel(int which, const PlainArgs& a, int ntiles, int threads, hipStream_t s) {
  switch (a.md.kind) {
    case MODEL_LG:
      if (a.md.dx == 1 && a.md.dy == 1) return plain_kernel_t<LG1>(which, a, ntiles, threads, s);
      if (a.md.dx == 2 && a.md.dy == 2) return plain_kernel_t<LG2>(which, a, ntiles, threads, s);
      return plain_kernel_t<LG4>(which, a, ntiles, threads, s);
    case MODEL_SV: return plain_kernel_t<SV>(which, a, ntiles, threads, s);
    case MODEL_SEIR: return plain_kernel_t<SEIR>(which, a, ntiles, threads, s);
    case MODEL_SPX: return plain_kernel_t<SPX>(which, a, ntiles, threads, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace mp
