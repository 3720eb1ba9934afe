// Plain form: one filter at a time, its particles spread over the whole GPU,
// two launches per observation (the shape of the reference's per-step XLA
// loop body, pfilter.py:97-182):
//
//  k_plain_prop  -- finishes the previous step from its scan tiles' totals
//                   (log-likelihood term, resampling decision), then the
//                   ancestor search (systematic comb against the previous
//                   step's prefix sums, staged per block in LDS) fused with the
//                   gather, the Euler sub-steps, the measurement density and the
//                   log-weight; per-block max / min.
//  k_plain_scan  -- fixed-point weights and their inclusive prefix sums within
//                   each scan tile, plus the tile's total.  Tiles do not wait
//                   on each other: the prefix of the tiles before a position is
//                   formed where it is needed, by the next propagate launch
//                   (each block from the tile totals, in LDS).  Every tile
//                   reduces the blocks' max log-weights (tile 0 also the
//                   minima and publishes both for the next launches).
//
// Every value a kernel reads was written by an earlier launch; no workgroup
// waits on another.
#include "blockops.hpp"
#include "engine.hpp"
#include "mech.hpp"

namespace mp {

constexpr int kMaxScanRows = 16;  // particles per scan thread, at most

// Chunk of propagate block b when a scan tile holds q chunks: blocks go to the
// kXcds XCDs round-robin, so within each group of kXcds * q blocks, block r
// (on XCD r % kXcds) takes chunk q * (r % kXcds) + r / kXcds; the q chunks of
// scan tile k of the group then all run on XCD k % kXcds, the scan tile's own.
__device__ __forceinline__ int xcd_chunk(int b, int q) {
  const int G = kXcds * q, g = b / G, r = b - g * G;
  return g * G + q * (r % kXcds) + r / kXcds;
}

// A wave walks the scan tiles' totals in chunks of 256 tiles: lane l holds
// tiles c0 + 4 l .. c0 + 4 l + 3 of the chunk at c0, its four loads in flight
// together, and a wave scan of the lane sums orders the chunk.
constexpr int kLaneTiles = 4, kChunkTiles = kLaneTiles * WAVE;

__device__ __forceinline__ void load_chunk(const PlainArgs& a, int c0, uint64_t (&v)[kLaneTiles]) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int i = 0; i < kLaneTiles; ++i) {
    const int k = c0 + kLaneTiles * lane + i;
    v[i] = k < a.ntiles ? a.agg[k] : 0;
  }
}

// Sum of the scan tiles' fixed-point totals of a step (the exact total of its
// weights), by one wave, given the first chunk's values `v` (load_chunk at 0).
__device__ uint64_t wave_step_total(const PlainArgs& a, const uint64_t (&v)[kLaneTiles]) {
  if (a.ends) return a.ends[a.ntiles - 1];
  uint64_t tot = lane_u64(wave_incl_scan_u64(v[0] + v[1] + v[2] + v[3]), WAVE - 1);
  for (int c0 = kChunkTiles; c0 < a.ntiles; c0 += kChunkTiles) {
    uint64_t w[kLaneTiles];
    load_chunk(a, c0, w);
    tot += lane_u64(wave_incl_scan_u64(w[0] + w[1] + w[2] + w[3]), WAVE - 1);
  }
  return tot;
}
__device__ uint64_t wave_step_total(const PlainArgs& a) {
  uint64_t v[kLaneTiles] = {};
  if (!a.ends) load_chunk(a, 0, v);
  return wave_step_total(a, v);
}

// What the end of step t fixes (the reference's _normalize_weights and the
// resampling criterion, helpers.py:34-57, pfilter.py:136-159), from the exact
// total of the step's fixed-point weights and its max / min log-weight.
struct PlainFin {
  double llt;      // log-likelihood term (float64)
  float llf;       // the float32 term (the normaliser of the weights)
  int collapsed;   // every weight was -inf
  int res;         // the step resamples
};

// plain_finish for a block that needs only the decision and, when the step
// does not resample, the normaliser: with no threshold (log threshold -inf) a
// step resamples unless its max is +inf or NaN (plain_finish's difference is
// then NaN), in which case step_loglik is NaN; a collapsed step (max -inf)
// resamples.  The same fields as plain_finish, without its float64 log.
__device__ PlainFin plain_finish_unthresholded(const PlainArgs& a, int t) {
  const float m = a.mm[2 * (t & 1)];
  PlainFin f;
  f.collapsed = (m == -__builtin_inff());
  f.res = !(m == __builtin_inff() || m != m);
  f.llt = f.res ? 0.0 : __builtin_nan("");  // unused when the step resamples
  f.llf = (float)f.llt;
  return f;
}

__device__ PlainFin plain_finish(const PlainArgs& a, int t, uint64_t total) {
  const int par = t & 1;
  const float m = a.mm[2 * par], mn = a.mm[2 * par + 1];
  const float logJ = a.logJ;
  PlainFin f;
  f.collapsed = (m == -__builtin_inff());
  f.llt = step_loglik(m, total, a.sh);
  f.llf = (float)f.llt;
  const float nmax = f.collapsed ? -logJ : m - f.llf;
  const float nmin = f.collapsed ? -logJ : mn - f.llf;
  f.res = (nmax - nmin) > a.logthresh;
  return f;
}

// Inclusive ends `e` of the lane's four tiles of a chunk, after `carry` (the
// end of the chunks before it), and the chunk's total: one DPP wave scan.
__device__ __forceinline__ uint64_t chunk_ends(const uint64_t (&v)[kLaneTiles], uint64_t carry,
                                               uint64_t (&e)[kLaneTiles]) {
  const uint64_t ls = v[0] + v[1] + v[2] + v[3];
  const uint64_t inc = wave_incl_scan_u64(ls);
  uint64_t x = carry + inc - ls;
#pragma unroll
  for (int i = 0; i < kLaneTiles; ++i) {
    x += v[i];
    e[i] = x;
  }
  return lane_u64(inc, WAVE - 1);
}

// #{i : C_i <= T} over the whole filter (its upper bound), by one wave, for
// T < total.  C_i is the prefix within scan tile k = i / L plus the end of the
// tiles before k.  With the ends in global memory a wave search finds k;
// otherwise the wave forms the ends from the tile totals, chunk by chunk (`v0`
// holds the first chunk): the ends <= T are the first k of them (they are
// non-decreasing), counted by ballots, and the base of tile k is end k - 1.
// Tiles past the last have total 0, so their ends equal the total and exceed T.
__device__ int64_t wave_global_upper_bound(const PlainArgs& a, const uint64_t (&v0)[kLaneTiles], uint64_t T) {
  int k;
  uint64_t base;
  if (a.ends) {
    k = (int)wave_upper_bound_u64(a.ends, 0, a.ntiles, T);
    base = k > 0 ? a.ends[k - 1] : 0;
  } else {
    uint64_t carry = 0;
    for (int c0 = 0;; c0 += kChunkTiles) {
      uint64_t v[kLaneTiles], e[kLaneTiles];
      if (c0 == 0) {
#pragma unroll
        for (int i = 0; i < kLaneTiles; ++i) v[i] = v0[i];
      } else {
        load_chunk(a, c0, v);
      }
      const uint64_t ctot = chunk_ends(v, carry, e);
      int n = 0;
#pragma unroll
      for (int i = 0; i < kLaneTiles; ++i) n += __popcll(__ballot(e[i] <= T));
      k = c0 + n;
      if (n < kChunkTiles) {
        base = carry;
        if (n > 0) {
          const int l = (n - 1) / kLaneTiles, i = (n - 1) % kLaneTiles;
          base = lane_u64(i == 0 ? e[0] : i == 1 ? e[1] : i == 2 ? e[2] : e[3], l);
        }
        break;
      }
      carry += ctot;
    }
  }
  // T < total, so tile k exists and its last prefix (its end) exceeds T
  const int64_t i0 = (int64_t)k << a.tile_shift, n = min((int64_t)a.tile_len, (int64_t)a.J - i0);
  return i0 + wave_upper_bound_u64(a.C + i0, 0, n, T - base);
}

// wave_global_upper_bound for a step whose tiles fit one chunk, from the
// tile ends `e` (chunk_ends of the first chunk with carry 0).
__device__ int64_t wave_upper_bound_one_chunk(const PlainArgs& a, const uint64_t (&e)[kLaneTiles], uint64_t T) {
  int n = 0;
#pragma unroll
  for (int i = 0; i < kLaneTiles; ++i) n += __popcll(__ballot(e[i] <= T));
  uint64_t base = 0;
  if (n > 0) {
    const int l = (n - 1) / kLaneTiles, i = (n - 1) % kLaneTiles;
    base = lane_u64(i == 0 ? e[0] : i == 1 ? e[1] : i == 2 ? e[2] : e[3], l);
  }
  const int64_t i0 = (int64_t)n << a.tile_shift, m = min((int64_t)a.tile_len, (int64_t)a.J - i0);
  return i0 + wave_upper_bound_u64(a.C + i0, 0, m, T - base);
}

// The prefix sum of particle i over the whole filter: the within-tile prefix
// plus the end of the tiles before its tile (`ends` in LDS or global memory).
__device__ __forceinline__ uint64_t global_prefix(const PlainArgs& a, const uint64_t* ends, int64_t i) {
  const int k = (int)(i >> a.tile_shift);
  return (k > 0 ? ends[k - 1] : 0) + a.C[i];
}

template <class M>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_plain_prop(const PlainArgs a) {
  __shared__ double red[32];
  __shared__ int64_t s_win[2];
  // Dynamic LDS: the ancestor window (a.win entries), then, when the tile ends
  // are not in global memory, the a.ntiles (<= kMaxLdsTiles) tile ends.  Sized
  // at launch, so that a filter of a few hundred tiles leaves room for eight
  // 256-thread blocks per CU.
  extern __shared__ uint64_t s_C[];
  uint64_t* const s_end = s_C + a.win;
  __shared__ uint64_t s_total;
  __shared__ PlainFin s_fin;
  const ModelDesc md = a.md;
  const int J = a.J, S = M::FIXED_S ? M::MAXS : md.S, t = a.t, tid = threadIdx.x, B = blockDim.x;
  // this step's scalars (row, grid, keys) and the previous step's comb uniform
  const PlainStep st = a.steps[t];
  const int ti0 = st.ti0;
  const int nst = st.nst;
  const float* dts = a.ser.dt_ext + a.ser.dt_off[a.unit];
  const float ty = st.ty;
  const float tstart = st.tstart;
  // the observation row, staged in LDS (published by the first barrier below)
  __shared__ float s_y[kSimMaxDy];
  if (tid < md.dy && tid < kSimMaxDy) s_y[tid] = a.ser.ys[st.obs * md.dy + tid];
  const float* y = s_y;
  const int64_t base = st.base;
  const float logJ = a.logJ;

  // Waves 0 to 2 (and 3 when the tile ends span several chunks) finish the
  // previous step below (a latency-bound chain of dependent loads): they issue
  // at a raised priority, so that the other waves' VALU work on the CU does
  // not delay it, and their first loads (the first chunk of scan tile totals,
  // block 0's running log-likelihood) are in flight during the parameter
  // staging.
  const int wv = tid >> 6, lane = tid & 63;
  const bool one_chunk = !a.ends && a.ntiles <= kChunkTiles;
  const bool pro = t > 0 && (wv < 3 || (wv == 3 && !a.ends && !one_chunk));
  if (pro) __builtin_amdgcn_s_setprio(2);
  uint64_t v0[kLaneTiles] = {};
  if (pro && !a.ends) load_chunk(a, 0, v0);
  const double ll_prev = (pro && blockIdx.x == 0 && wv == 2 && t > 1) ? a.st->loglik : 0.0;

  // Parameters prepared once per filter (k_plain_prepare), staged in LDS.  The
  // first barrier below publishes them; a model whose noise draw reads them
  // waits for them here.
  __shared__ typename M::P s_P;
  if (tid == 0) s_P = *reinterpret_cast<const typename M::P*>(a.params);
  if constexpr (DrawUsesP<M>::value) __syncthreads();
  const typename M::P& P = s_P;

  const float* Xc = (t & 1) ? a.X1 : a.X0;
  float* Xn = (t & 1) ? a.X0 : a.X1;

  float lmax = -__builtin_inff(), lmin = __builtin_inff();
  const int chunk = a.xcd_q > 0 ? xcd_chunk(blockIdx.x, a.xcd_q) : (int)blockIdx.x;
  const int64_t tile0 = (int64_t)chunk * a.prop_len;
  const int64_t tile1 = min((int64_t)J, tile0 + (int64_t)a.prop_len);
  if (tile0 >= J) {  // a trailing block without particles (never block 0)
    if (tid == 0) {
      a.pmax[blockIdx.x] = lmax;
      a.pmin[blockIdx.x] = lmin;
    }
    return;
  }

  // End of the previous step, t - 1 (from its scan tiles' totals):
  //  * waves 0 and 1 find the two ends of this block's ancestor window:
  //    A(T) = #{i : C_i <= T} is non-decreasing in the comb position, so every
  //    j of the block has A(T_j) in [A(T_tile0), A(T_tile1-1)]; that slice of
  //    C is staged in LDS below;
  //  * wave 2 computes the step's log-likelihood term and resampling decision
  //    (block 0 also adds the term to the running log-likelihood);
  //  * wave 3 forms the tile ends in LDS for the staging.
  // The searches run whether or not the step resamples (wave 2 decides that
  // concurrently); their result is then unused.
  if (t > 0) {
    if (wv < 2 && one_chunk) {
      // the tile ends from one DPP scan; wave 0 also stores them in LDS
      uint64_t e[kLaneTiles];
      const uint64_t total = chunk_ends(v0, 0, e);
      if (wv == 0) {
#pragma unroll
        for (int i = 0; i < kLaneTiles; ++i) {
          const int k = kLaneTiles * lane + i;
          if (k < a.ntiles) s_end[k] = e[i];
        }
      }
      const uint64_t T = comb_threshold(st.ures_prev, wv == 0 ? tile0 : tile1 - 1, J, total);
      // every prefix sum is <= total
      const int64_t v = T >= total ? J : wave_upper_bound_one_chunk(a, e, T);
      if (lane == 0) s_win[wv] = v;
    } else if (wv < 3) {
      const uint64_t total = wave_step_total(a, v0);
      if (wv < 2) {
        const int64_t jj = wv == 0 ? tile0 : tile1 - 1;
        const uint64_t T = comb_threshold(st.ures_prev, jj, J, total);
        // every prefix sum is <= total
        const int64_t v = T >= total ? J : wave_global_upper_bound(a, v0, T);
        if (lane == 0) s_win[wv] = v;
      } else {
        // block 0 keeps the running log-likelihood and needs the term; with no
        // threshold the others need only the decision
        const PlainFin f = (blockIdx.x != 0 && a.logthresh == -__builtin_inff())
                               ? plain_finish_unthresholded(a, t - 1)
                               : plain_finish(a, t - 1, total);
        if (lane == 0) {
          s_fin = f;
          s_total = total;
          if (blockIdx.x == 0) {
            a.st->loglik = ll_prev + f.llt;
            if (a.cll) a.cll[t - 1] = f.llf;
          }
        }
      }
    } else if (wv == 3 && !a.ends && !one_chunk) {
      uint64_t carry = 0;
      for (int c0 = 0; c0 < a.ntiles; c0 += kChunkTiles) {
        uint64_t v[kLaneTiles], e[kLaneTiles];
        if (c0 == 0) {
#pragma unroll
          for (int i = 0; i < kLaneTiles; ++i) v[i] = v0[i];
        } else {
          load_chunk(a, c0, v);
        }
        carry += chunk_ends(v, carry, e);
#pragma unroll
        for (int i = 0; i < kLaneTiles; ++i) {
          const int k = c0 + kLaneTiles * lane + i;
          if (k < a.ntiles) s_end[k] = e[i];
        }
      }
    }
  }
  if (pro) __builtin_amdgcn_s_setprio(0);
  // Models whose step noise depends on the particle's key alone draw it for
  // the first kPreRows rows here: the other waves do this VALU work while
  // waves 0 to 3 wait on their loads (and all of them while the parameters
  // arrive, when the draw does not read them).
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
  __syncthreads();
  const bool have_anc = t > 0 && s_fin.res;
  const bool uniform_w = t == 0 || s_fin.res;
  const bool collapsed = t > 0 && s_fin.collapsed;
  const float llf = t > 0 ? s_fin.llf : 0.0f;
  const float wconst = t == 0 ? a.log1J : 0.0f - logJ;
  const uint64_t total = t > 0 ? s_total : 0;
  const uint64_t* ends = a.ends ? a.ends : s_end;
  const int64_t wlo = have_anc ? s_win[0] : 0, whi = have_anc ? s_win[1] : 0;
  const bool in_lds = have_anc && (whi - wlo) <= a.win;
  if (in_lds) {
    const int n = (int)(whi - wlo);
    if (a.ends) {
      for (int i = tid; i < n; i += B) s_C[i] = global_prefix(a, a.ends, wlo + i);
    } else {
      for (int i = tid; i < n; i += B) s_C[i] = global_prefix(a, s_end, wlo + i);
    }
    __syncthreads();
  }
  auto row = [&](int r, bool use_z, const float (&zp)[NZ]) {
    // A compiler barrier (no instruction): the model parameters are read from
    // LDS inside each row rather than held in registers across the rows, which
    // would spill at the 8-wave register budget.
    __asm__ volatile("" ::: "memory");
    const int64_t j = tile0 + (int64_t)r * B + tid;
    if (j >= tile1) return;
    float x[M::MAXS];
    int64_t src = j;
    if (t == 0) {
      m_rinit<M>(md, P, split_at(Key{a.fkey[0], a.fkey[1]}, 1 + (uint64_t)j), a.ser.t0, x,
                 cov_row(a.ser, a.unit, 0));
    } else {
      if (have_anc) {
        const uint64_t tj = comb_threshold(st.ures_prev, j, J, total);
        int64_t an;
        if (in_lds) {
          an = wlo + (whi > wlo ? upper_bound_fixed(s_C, (int)(whi - wlo), tj) : 0);
        } else {
          int64_t lo = wlo, hi = whi;
          while (lo < hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            if (global_prefix(a, ends, mid) <= tj) lo = mid + 1;
            else hi = mid;
          }
          an = lo;
        }
        src = an < J ? an : J - 1;
      }
      MP_UNROLL for (int s = 0; s < M::MAXS; ++s) if (s < S) x[s] = Xc[(int64_t)s * J + src];
    }
    float wprev = uniform_w ? wconst : (collapsed ? -logJ : a.lw[j] - llf);
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

template <class M>
__global__ __launch_bounds__(1024) void k_plain_scan(const PlainArgs a) {
  __shared__ double red[32];
  __shared__ uint64_t tab[17];  // wave totals (<= 16), then their exclusive prefix
  const int J = a.J, tid = threadIdx.x, B = blockDim.x, lane = tid & 63, w = tid >> 6;
  const int nw = (B + 63) >> 6;
  const int par = a.t & 1;
  const int sh = a.sh;
  const int tile = blockIdx.x;
  const int rows = a.rows;
  const int64_t tile0 = (int64_t)tile * a.tile_len;
  // Each thread owns `rows` consecutive particles j0 .. j0 + rows - 1: it
  // loads them as float4 vectors, scans them serially and takes one wave scan
  // of its total (rather than one wave scan per row).
  const int64_t j0 = tile0 + (int64_t)tid * rows;
  const bool full = j0 + rows <= J;  // all of the thread's particles exist
  const bool vec = full && (rows & 3) == 0;

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
  // the step's max (and, in tile 0, min) log-weight from the propagate
  // blocks' values, eight independent loads in flight per thread
  float m = -__builtin_inff(), mn = __builtin_inff();
  for (int i0 = 0; i0 < a.nprop; i0 += 8 * B) {
    float vx[8], vn[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + k * B + tid;
      vx[k] = i < a.nprop ? a.pmax[i] : -__builtin_inff();
      vn[k] = (tile == 0 && i < a.nprop) ? a.pmin[i] : __builtin_inff();
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      m = nanmax(m, vx[k]);
      mn = nanmin(mn, vn[k]);
    }
  }
  if (tile == 0) {
    block_maxmin(m, mn, red);
    if (tid == 0) {
      a.mm[2 * par] = m;
      a.mm[2 * par + 1] = mn;
    }
  } else {
    m = block_nanmax(m, red);
  }

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
  // wave 0: exclusive scan of the wave totals and the tile's total
  if (w == 0) {
    const uint64_t v = lane < nw ? tab[lane] : 0;
    const uint64_t inc = wave_incl_scan_u64(v);
    if (lane < nw) tab[lane] = inc - v;
    if (lane == nw - 1) a.agg[tile] = inc;
  }
  __syncthreads();
  const uint64_t base = tab[w] + texcl;
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
}

// Inclusive prefix of the scan tiles' totals into a.ends (one workgroup), for
// filters with more tiles than a propagate block forms in LDS: each thread
// sums a run of consecutive tiles, a block scan orders the runs, and a second
// pass writes the ends.
__global__ __launch_bounds__(1024) void k_plain_ends(const PlainArgs a) {
  __shared__ uint64_t tab[17];
  const int tid = threadIdx.x, B = blockDim.x, lane = tid & 63, w = tid >> 6, nw = (B + 63) >> 6;
  const int q = (a.ntiles + B - 1) / B;
  const int64_t k0 = (int64_t)tid * q;
  uint64_t v = 0;
  for (int i = 0; i < q; ++i)
    if (k0 + i < a.ntiles) v += a.agg[k0 + i];
  const uint64_t winc = wave_incl_scan_u64(v);
  if (lane == 63) tab[w] = winc;
  __syncthreads();
  if (w == 0) {
    const uint64_t x = lane < nw ? tab[lane] : 0;
    const uint64_t inc = wave_incl_scan_u64(x);
    if (lane < nw) tab[lane] = inc - x;
  }
  __syncthreads();
  uint64_t e = tab[w] + winc - v;
  for (int i = 0; i < q; ++i)
    if (k0 + i < a.ntiles) {
      e += a.agg[k0 + i];
      a.ends[k0 + i] = e;
    }
}

// After the last step: its log-likelihood term completes the filter's
// log-likelihood (one wave).
__global__ __launch_bounds__(64) void k_plain_final(const PlainArgs a) {
  if (a.T == 0) {
    if (threadIdx.x == 0) *a.loglik_out = 0.0;
    return;
  }
  const uint64_t total = wave_step_total(a);
  const PlainFin f = plain_finish(a, a.T - 1, total);
  if (threadIdx.x == 0) {
    *a.loglik_out = (a.T == 1 ? 0.0 : a.st->loglik) + f.llt;
    if (a.cll) a.cll[a.T - 1] = f.llf;
  }
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
  __shared__ PlainFin s_fin;
  if (tid < WAVE) {
    const PlainFin f = plain_finish(a, a.t, wave_step_total(a));
    if (tid == 0) s_fin = f;
  }
  __syncthreads();
  const PlainFin st = s_fin;
  const float logJ = a.logJ;
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
  static_assert(sizeof(typename M::P) <= kParamSlot, "prepared parameters exceed their slot");
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
      if (md.dx <= 4 && md.dy <= 4) return plain_prepare_t<LG4>(md, theta, stride, R, flags, out, pstride, s);
      return plain_prepare_t<LG8>(md, theta, stride, R, flags, out, pstride, s);
    case MODEL_SV: return plain_prepare_t<SV>(md, theta, stride, R, flags, out, pstride, s);
    case MODEL_SEIR: return plain_prepare_t<SEIR>(md, theta, stride, R, flags, out, pstride, s);
    case MODEL_SPX: return plain_prepare_t<SPX>(md, theta, stride, R, flags, out, pstride, s);
  }
  return hipErrorInvalidValue;
}

// Start of a filter: the step table (one thread per observation) and the
// running state.
__global__ void k_plain_reset(const PlainArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.T) {
    const int64_t obs = a.ser.obs_off[a.unit] + t;
    const int ti0 = a.ser.tidx0[obs], nst = a.ser.nstep[obs];
    const int64_t base = ti0 + (int64_t)t;
    PlainStep s;
    s.obs = obs;
    s.base = base;
    s.ti0 = ti0;
    s.nst = nst;
    s.tstart = t == 0 ? a.ser.t0 : a.ser.times[obs - 1];
    s.ty = a.ser.times[obs];
    s.dt0 = nst > 0 ? a.ser.dt_ext[a.ser.dt_off[a.unit] + ti0] : 0.0f;
    s.k0 = a.sched[2 * base];
    s.k1 = a.sched[2 * base + 1];
    s.ures_prev = 0.0f;
    if (t > 0) {
      // step t - 1 resamples with the key after its sub-step keys
      const int64_t rb = a.ser.tidx0[obs - 1] + (int64_t)(t - 1) + a.ser.nstep[obs - 1];
      s.ures_prev = uniform01(bits_at(Key{a.sched[2 * rb], a.sched[2 * rb + 1]}, 0));
    }
    a.steps[t] = s;
  }
  if (t == 0) a.st->loglik = 0.0;
}

hipError_t plain_reset(const PlainArgs& a, hipStream_t s) {
  const int n = a.T > 0 ? a.T : 1;
  mp_launch(k_plain_reset, dim3((n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

// which = 0: propagate kernel, 1: scan kernel (and the tile ends when they are
// global), 2: step diagnostics, 3: the filter's last step
template <class M>
static hipError_t plain_kernel_t(int which, const PlainArgs& a, int ntiles, int threads, hipStream_t s) {
  if (which == 0) {
    const int64_t lds_words = (int64_t)a.win + (a.ends ? 0 : a.ntiles);
    mp_launch(k_plain_prop<M>, dim3(a.nprop), dim3(threads), (uint32_t)(lds_words * sizeof(uint64_t)), s, a);
  } else if (which == 1) {
    mp_launch(k_plain_scan<M>, dim3(ntiles), dim3(threads), 0, s, a);
    if (a.ends) mp_launch(k_plain_ends, dim3(1), dim3(1024), 0, s, a);
  } else if (which == 2) {
    mp_launch(k_plain_diag<M>, dim3(a.ndiag), dim3(256), 0, s, a);
    mp_launch(k_plain_diag_final<M>, dim3(1), dim3(64), 0, s, a);
  } else {
    mp_launch(k_plain_final, dim3(1), dim3(64), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t plain_kernel(int which, const PlainArgs& a, int ntiles, int threads, hipStream_t s) {
  switch (a.md.kind) {
    case MODEL_LG:
      if (a.md.dx == 1 && a.md.dy == 1) return plain_kernel_t<LG1>(which, a, ntiles, threads, s);
      if (a.md.dx == 2 && a.md.dy == 2) return plain_kernel_t<LG2>(which, a, ntiles, threads, s);
      if (a.md.dx <= 4 && a.md.dy <= 4) return plain_kernel_t<LG4>(which, a, ntiles, threads, s);
      return plain_kernel_t<LG8>(which, a, ntiles, threads, s);
    case MODEL_SV: return plain_kernel_t<SV>(which, a, ntiles, threads, s);
    case MODEL_SEIR: return plain_kernel_t<SEIR>(which, a, ntiles, threads, s);
    case MODEL_SPX: return plain_kernel_t<SPX>(which, a, ntiles, threads, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace mp
