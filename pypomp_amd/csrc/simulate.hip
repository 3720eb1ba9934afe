// Simulation of states and observations: _simulate_internal / _simulate_helper
// (pypomp/core/algorithms/simulate.py:15-142), vmapped over parameter rows
// (simulate.py:71-74).  One thread per (row, simulation); every thread walks
// its row's key chain itself (a few threefry calls per observation):
//   split(key, nsim+1): [0] running key, [1+s] rinit key of simulation s;
//   per observation i: (key, sub) = split(key); Euler sub-steps
//   (next, sk) = split(sub), simulation key split(sk, nsim)[s], sub = next
//   (model_mechanics.py:894-906, 584-588); then split(key, nsim+1): [0] the
//   running key, [1+s] the measurement key of simulation s.
// Times follow the reference's loop exactly: the process of interval i starts
// at t = t0 for i <= 1 and at times[i-2] after (the loop carries t = times0[i]
// set after the step, simulate.py:119-131), and rmeas sees times0[i].
// Outputs in the reference's internal layout, coalesced over simulations:
// X (R, T+1, S, nsim), Y (R, T, dy, nsim).
#include "engine.hpp"
#include "mech.hpp"

namespace mp {

namespace {

__device__ __forceinline__ Key sim_key(const uint32_t* p, int64_t i) { return Key{p[2 * i], p[2 * i + 1]}; }

template <class M>
__global__ __launch_bounds__(256) void k_simulate(const SimArgs a) {
  const int r = blockIdx.y;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nsim) return;
  const ModelDesc md = a.md;
  const int S = M::FIXED_S ? M::MAXS : md.S, d = md.d, dy = md.dy, T = a.ser.Tmax;
  const int64_t ns = a.nsim;
  typename M::P P;
  {
    float th[M::MAXP];
    MP_UNROLL for (int p = 0; p < M::MAXP; ++p) if (p < d) th[p] = a.theta[(int64_t)r * a.theta_stride + p];
    if ((a.flags & F_SHOULD_TRANS) && !md.identity_trans) M::from_est(md, th);
    M::prepare(md, th, P);
  }
  const Key k0 = sim_key(a.keys, r);
  Key run = split_at(k0, 0);
  float x[M::MAXS];
  m_rinit<M>(md, P, split_at(k0, 1 + (uint64_t)s), a.ser.t0, x, cov_row(a.ser, 0, 0));
  float* Xr = a.X + (int64_t)r * (T + 1) * S * ns;
  float* Yr = a.Y + (int64_t)r * T * dy * ns;
  MP_UNROLL for (int c = 0; c < M::MAXS; ++c) if (c < S) Xr[c * ns + s] = x[c];
  float t = a.ser.t0;
  for (int i = 0; i < T; ++i) {
    Key sub = split_at(run, 1);
    run = split_at(run, 0);
    MP_UNROLL for (int c = 0; c < M::MAXS; ++c) if ((md.accum_mask >> c) & 1u) x[c] = 0.f;
    const int nst = a.ser.nstep[i], ti0 = a.ser.tidx0[i];
    float tt = t;
    for (int k = 0; k < nst; ++k) {
      const Key nxt = split_at(sub, 0), sk = split_at(sub, 1);
      const float dt = a.ser.dt_ext[ti0 + k];
      m_step<M>(md, P, sk, (uint64_t)s, tt, dt, x, cov_row(a.ser, 0, ti0 + k));
      tt = tt + dt;
      sub = nxt;
    }
    t = i == 0 ? a.ser.t0 : a.times[i - 1];
    const Key mk = split_at(run, 1 + (uint64_t)s);
    run = split_at(run, 0);
    float y[kSimMaxDy];
    m_rmeas<M>(md, P, mk, x, t, y, cov_row(a.ser, 0, ti0 + nst));
    float* Xo = Xr + (int64_t)(i + 1) * S * ns;
    MP_UNROLL for (int c = 0; c < M::MAXS; ++c) if (c < S) Xo[c * ns + s] = x[c];
    float* Yo = Yr + (int64_t)i * dy * ns;
    for (int c = 0; c < dy; ++c) Yo[c * ns + s] = y[c];
  }
}

template <class M>
hipError_t launch_simulate_t(const SimArgs& a, hipStream_t st) {
  const int block = a.nsim >= 256 ? 256 : ((a.nsim + 63) / 64) * 64;
  const dim3 grid((unsigned)((a.nsim + block - 1) / block), (unsigned)a.R);
  hipLaunchKernelGGL((k_simulate<M>), grid, dim3(block), 0, st, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_simulate(const SimArgs& a, hipStream_t st) {
  switch (a.md.kind) {
    case MODEL_LG:
      if (a.md.dx == 1 && a.md.dy == 1) return launch_simulate_t<LG1>(a, st);
      if (a.md.dx == 2 && a.md.dy == 2) return launch_simulate_t<LG2>(a, st);
      if (a.md.dx <= 4 && a.md.dy <= 4) return launch_simulate_t<LG4>(a, st);
      return launch_simulate_t<LG8>(a, st);
    case MODEL_SV: return launch_simulate_t<SV>(a, st);
    case MODEL_SEIR: return launch_simulate_t<SEIR>(a, st);
    case MODEL_SPX: return launch_simulate_t<SPX>(a, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace mp
