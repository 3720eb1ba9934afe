// How the kernels call a model's plugins (the reference's _RInit / _RProc /
// _DMeas / _RMeas wrappers, pypomp/core/model_mechanics.py:361-760).  A model
// with EXT (spx.hpp) reads covariates and defines the @vectorized draw; the
// others take neither, and their plugins are called as before.
//
// Keys of one Euler sub-step (model_mechanics.py:542-612): the default rproc
// gives particle j the key split(sk, J)[j] (step_pf, :584-588); a @vectorized
// rproc gets the sub-step key sk itself and draws for all particles at once
// (step_vec, :553-568), so particle j's variate is element j of that draw.
#pragma once

#include <type_traits>

#include "seir.hpp"
#include "spx.hpp"

namespace mp {

template <class M, class = void>
struct IsExt : std::false_type {};
template <class M>
struct IsExt<M, std::void_t<decltype(M::EXT)>> : std::integral_constant<bool, M::EXT> {};

template <class M>
__device__ __forceinline__ void m_rinit(const ModelDesc& md, const typename M::P& p, Key k, float t0, float* x,
                                        const float* cv) {
  if constexpr (IsExt<M>::value) M::rinit_x(md, p, k, t0, x, cv);
  else M::rinit(md, p, k, t0, x);
}

// One Euler sub-step of particle j from the sub-step key sk.
template <class M>
__device__ __forceinline__ void m_step(const ModelDesc& md, const typename M::P& p, Key sk, uint64_t j, float t,
                                       float dt, float* x, const float* cv) {
  if constexpr (IsExt<M>::value) {
    if (md.vectorized) M::step_x(md, p, sk, j, t, dt, x, cv);
    else M::step_x(md, p, split_at(sk, j), 0, t, dt, x, cv);
  } else {
    M::step(md, p, split_at(sk, j), t, dt, x);
  }
}

// Whether a model's draw reads its prepared parameters (a model says so with
// DRAW_USES_P = false; otherwise it is taken to read them).
template <class M, class = void>
struct DrawUsesP : std::true_type {};
template <class M>
struct DrawUsesP<M, std::void_t<decltype(M::DRAW_USES_P)>> : std::integral_constant<bool, M::DRAW_USES_P> {};

// The sub-step noise of particle j (models whose noise depends on the key
// only, NOISE > 0): m_draw + m_step_z == m_step.
template <class M>
__device__ __forceinline__ void m_draw(const ModelDesc& md, const typename M::P& p, Key sk, uint64_t j, float* z) {
  if constexpr (IsExt<M>::value) {
    if (md.vectorized) M::draw_x(p, sk, j, z);
    else M::draw_x(p, split_at(sk, j), 0, z);
  } else {
    M::draw(p, split_at(sk, j), z);
  }
}

template <class M>
__device__ __forceinline__ void m_step_z(const ModelDesc& md, const typename M::P& p, const float* z, float t,
                                         float dt, float* x, const float* cv) {
  if constexpr (IsExt<M>::value) M::step_z_x(md, p, z, t, dt, x, cv);
  else M::step_z(md, p, z, t, dt, x);
}

template <class M>
__device__ __forceinline__ float m_dmeas(const ModelDesc& md, const typename M::P& p, const float* y,
                                         const float* x, float t, const float* cv) {
  if constexpr (IsExt<M>::value) return M::dmeas_x(md, p, y, x, t, cv);
  else return M::dmeas(md, p, y, x, t);
}

template <class M>
__device__ __forceinline__ void m_rmeas(const ModelDesc& md, const typename M::P& p, Key k, const float* x, float t,
                                        float* y, const float* cv) {
  if constexpr (IsExt<M>::value) M::rmeas_x(md, p, k, x, t, y, cv);
  else M::rmeas(md, p, k, x, t, y);
}

}  // namespace mp
