"""GPU parity of the MOP Hessian (mp_mop_hess) against the oracle: the value
reproduces mp_mop bit for bit (same ancestry), the gradient the autograd
gradient, and the Hessian torch's double backward of the frozen-ancestry
objective -- the restatement of jax.hessian of _mop_internal_mean
(core/algorithms/ad_helpers.py:8, 56-75)."""

import numpy as np
import pytest

from oracle import lg as olg, mop_grad as mg, sv as osv, threefry as tf
from tests.helpers import workloads

pytestmark = pytest.mark.gpu
F32 = np.float32

# float32 value parts on the GPU, float64 throughout in the oracle
HESS_RTOL = 1e-3


def sv_struct(T):
    import pypomp_amd.functional as F
    from pypomp_amd.models import SVModel
    from pypomp_amd.series import Series
    ys = workloads()["sv_seed4_ys"][:T]
    return F.PompStruct(model=SVModel(), units=[Series.build(ys, np.arange(1, T + 1.0), 0.0, nstep=1)]), ys


def lg2_struct(T):
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    w = workloads()
    ys = w["lg2_seed1_ys"][:T]
    return (F.PompStruct(model=LGModel(2, 2, separate_init=True),
                         units=[Series.build(ys, np.arange(1, T + 1.0), 0.0, nstep=1)]), ys, w["lg2_seed1_theta"])


def close(got, exp, rtol):
    scale = np.abs(exp).max()
    np.testing.assert_allclose(got, exp, rtol=rtol, atol=rtol * scale)


@pytest.mark.parametrize("alpha", [0.97, 0.5, 0.0])
def test_sv_hessian_matches_oracle(alpha):
    import pypomp_amd.functional as F
    T, J = 40, 64
    struct, ys = sv_struct(T)
    th = osv.SVModel().to_est(osv.default_theta())
    key = tf.split(tf.key(9), 1)[0]
    print(f"sv hessian: alpha={alpha} T={T} J={J}", flush=True)
    ll, g, H = F.mop_hess(struct, th[None], J, alpha, key[None])
    # same key tree, same ancestry: the value of mp_mop, bit for bit
    assert -F.mop(struct, th[None], J, alpha, key[None], return_float64=True)[0] == ll[0]
    val, og, oH, _ = mg.hessian_and_fd(mg.sv_hessian_case(th, ys, key, J, alpha), th, T)
    close(g[0], og, HESS_RTOL)
    np.testing.assert_array_equal(H[0], H[0].T)
    close(H[0], oH, HESS_RTOL)


def test_lg2_separate_init_hessian_matches_oracle():
    import pypomp_amd.functional as F
    T, J = 30, 48
    struct, ys, theta = lg2_struct(T)
    model = olg.LGModel(olg.LGSpec(2, 2, separate_init=True))
    th = model.to_est(theta)
    key = tf.split(tf.key(10), 1)[0]
    ll, g, H = F.mop_hess(struct, th[None], J, 0.9, key[None])
    assert -F.mop(struct, th[None], J, 0.9, key[None], return_float64=True)[0] == ll[0]
    val, og, oH, _ = mg.hessian_and_fd(mg.lg_hessian_case(model, th, ys, key, J, 0.9), th, T)
    assert H.shape == (1, 19, 19)
    close(g[0], og, HESS_RTOL)
    close(H[0], oH, HESS_RTOL)


def test_lg1_hessian_matches_oracle():
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    from tests.helpers import lg1_data, lg1_theta
    ys = np.concatenate([lg1_data(0), lg1_data(1), lg1_data(2)])
    struct = F.PompStruct(model=LGModel(1, 1), units=[Series.build(ys, np.arange(1, 13.0), 0.0, nstep=1)])
    model = olg.LGModel(olg.LGSpec(1, 1))
    th = model.to_est(lg1_theta())
    key = tf.split(tf.key(12), 1)[0]
    ll, g, H = F.mop_hess(struct, th[None], 100, 0.8, key[None])
    val, og, oH, _ = mg.hessian_and_fd(mg.lg_hessian_case(model, th, ys, key, 100, 0.8), th, 12)
    close(g[0], og, HESS_RTOL)
    close(H[0], oH, HESS_RTOL)


def test_rows_are_independent_and_the_gradient_matches_the_adjoint():
    import pypomp_amd.functional as F
    T, J = 60, 200
    struct, ys = sv_struct(T)
    base = osv.SVModel().to_est(osv.default_theta())
    rng = np.random.default_rng(1)
    th = (base[None] + F32(0.05) * rng.standard_normal((3, 6)).astype(F32)).astype(F32)
    keys = tf.split(tf.key(13), 3)
    ll, g, H = F.mop_hess(struct, th, J, 0.97, keys)
    for r in range(3):
        l1, g1, H1 = F.mop_hess(struct, th[r:r + 1], J, 0.97, keys[r:r + 1])
        assert l1[0] == ll[r]
        np.testing.assert_array_equal(g1[0], g[r])
        np.testing.assert_array_equal(H1[0], H[r])
    # the reverse-mode adjoint's gradient of the same objective
    lg, gg = F.mop_grad(struct, th, J, 0.97, keys, K=10)
    np.testing.assert_allclose(lg, ll, rtol=1e-12)
    # the adjoint's bar against the autograd gradient (tests/test_gpu_mop_grad.py)
    close(g, gg, 1e-2)


def test_unsupported_model_raises():
    import pypomp_amd.functional as F
    from pypomp_amd import _native
    from pypomp_amd.models import SEIRModel
    from pypomp_amd.series import Series
    ys = workloads()["seir_seed3_ys"][:10]
    struct = F.PompStruct(model=SEIRModel(), units=[Series.build(ys, np.arange(1, 11.0), 0.0, nstep=7)])
    th = np.zeros((1, 10), F32)
    with pytest.raises(_native.NativeError, match="no Hessian"):
        F.mop_hess(struct, th, 100, 0.9, tf.split(tf.key(1), 1))
