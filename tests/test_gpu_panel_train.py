"""GPU parity of panel MOP gradient ascent (functional.panel_train) against the
reference's panel-train baseline (tests/test_regression/
test_panel_train_regression.py) and the oracle's restatement of
core/algorithms/train.py:190-415 driven by its own autograd MOP gradients:
chunked units, shared parameters carried across chunks, the running chunk
count for the shared optimizer and the iteration count for the unit one."""

import numpy as np
import pytest

from oracle import filters as fl
from oracle import lg as olg, mop_grad as mg, sv as osv, threefry as tf
from tests.helpers import assert_close_tol, lg1_data, lg1_theta, workloads

pytestmark = pytest.mark.gpu
F32 = np.float32
LG_NAMES = ["A11", "C11", "Q11", "R11", "X0_1"]


def lg_long(T, seed):
    model = olg.LGModel(olg.LGSpec(1, 1))
    ser = fl.make_series(np.zeros((T, 1)), np.arange(1, T + 1.0), 0.0, nstep=1)
    _, Y = fl.simulate(model, lg1_theta(), ser, tf.split(tf.key(seed), 1)[0])
    return Y[:, 0, :].astype(F32)


def panel_struct(model, data, shared, unit, par_trans=None):
    import pypomp_amd.functional as F
    from pypomp_amd.series import Series
    units = [Series.build(y, np.arange(1, len(y) + 1.0), 0.0, nstep=1) for y in data]
    return F.PanelStruct(model=model, units=units, shared_param_names=shared, unit_param_names=unit,
                         par_trans=par_trans)


def test_panel_train_regression(golden):
    # the reference's fixture: units on data seeds 0 and 1, shared A11 and
    # C11, three Adam steps, eta 0.01, alpha 0.97, chunk_size 1, keys
    # split(key(20260811), n M U); identity panel transform
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    struct = panel_struct(LGModel(1, 1), [lg1_data(0), lg1_data(1)], ["A11", "C11"], ["Q11", "R11", "X0_1"],
                          par_trans=pp.ParTrans())
    th = lg1_theta()
    eta = pp.LearningRate({n: 0.01 for n in LG_NAMES})
    keys = tf.split(tf.key(20260811), 6).reshape(1, 3, 2, 2)
    nll, sh, un = F.panel_train(struct, th[None, :2], np.stack([th[2:], th[2:]])[None], 12, pp.Adam(), 3, eta,
                                0.97, keys, alpha_cooling=1.0, chunk_size=1)
    g = golden["test_panel_train_regression"]
    assert nll.shape == (1, 4) and sh.shape == (1, 4, 2) and un.shape == (1, 4, 2, 3)
    assert np.isnan(nll[0, 0])
    assert_close_tol(nll[:, -1], g["final_neg_loglik"])
    assert_close_tol(sh[:, -1], g["final_shared"])
    assert_close_tol(un[:, -1], g["final_unit"])


CASES = [
    # (optimizer, chunk_size, shared names, unit names)
    ("adam", 1, ["A11", "C11"], ["Q11", "R11", "X0_1"]),
    ("adam", 2, ["C11", "A11"], ["X0_1", "R11", "Q11"]),
    ("sgd", 4, ["A11"], ["C11", "Q11", "R11", "X0_1"]),
    ("fmadam", 2, ["A11", "Q11"], ["C11", "R11", "X0_1"]),
    ("adam", 1, [], LG_NAMES),
]


@pytest.mark.parametrize("opt,cs,shared,unit", CASES)
def test_lg_panel_train_matches_oracle(opt, cs, shared, unit):
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    U, T, J, M, n = 4, 30, 100, 5, 2
    data = [lg_long(T, 40 + u) for u in range(U)]
    struct = panel_struct(LGModel(1, 1), data, shared, unit)
    omodel = olg.LGModel(olg.LGSpec(1, 1))
    names = shared + unit
    perm = np.array([names.index(p) for p in LG_NAMES])
    rng = np.random.default_rng(5)
    base = lg1_theta()
    nat = base[None, None] * np.exp(F32(0.1) * rng.standard_normal((n, U, 5))).astype(F32)
    nat[..., 4] = F32(0.2) * rng.standard_normal((n, U)).astype(F32)       # X0_1 may be negative
    nat = nat.astype(F32)
    idx = [LG_NAMES.index(p) for p in names]
    sh0 = nat[:, 0, idx[:len(shared)]]
    un0 = nat[:, :, idx[len(shared):]]
    rates = np.linspace(0.03, 0.01, M)
    eta = pp.LearningRate({p: rates for p in LG_NAMES})
    keys = tf.split(tf.key(77), n * M * U).reshape(n, M, U, 2)
    optimizer = {"adam": pp.Adam(), "sgd": pp.SGD(), "fmadam": pp.FullMatrixAdam()}[opt]
    print(f"panel train: {opt} chunk {cs} shared {shared}", flush=True)
    nll, sh, un = F.panel_train(struct, sh0, un0, J, optimizer, M, eta, 0.9, keys, alpha_cooling=0.5,
                                chunk_size=cs)

    def vg(th, k, alpha, u):
        val, _, gr = mg.lg_mop_value_and_grad(omodel, th, data[u], k, J, alpha)
        return val, gr

    for r in range(n):
        # the model's transform maps the natural start to the estimation scale
        full = np.concatenate([np.broadcast_to(sh0[r][None], (U, len(shared))), un0[r]], axis=1)
        est = omodel.to_est(full[:, perm])[:, np.argsort(perm)].astype(F32)
        onll, osh, oun = mg.panel_train(vg, est[0, :len(shared)], est[:, len(shared):], perm, keys[r], T, M,
                                        eta.to_array(shared, M).reshape(M, len(shared)),
                                        eta.to_array(unit, M).reshape(M, len(unit)), 0.9, alpha_cooling=0.5,
                                        chunk_size=cs, optimizer=opt)
        full_o = np.concatenate([np.broadcast_to(osh[:, None, :], (M + 1, U, len(shared))), oun], axis=2)
        nat_o = omodel.from_est(full_o[..., perm])[..., np.argsort(perm)]
        np.testing.assert_allclose(nll[r, 1:], onll[1:], rtol=1e-4)
        np.testing.assert_allclose(sh[r], nat_o[:, 0, :len(shared)], rtol=1e-4, atol=1e-5)
        np.testing.assert_allclose(un[r], nat_o[..., len(shared):], rtol=1e-4, atol=1e-5)


def test_sv_panel_train_matches_oracle():
    # the stochastic-volatility adjoint under the panel loop: two units of
    # different data, shared mu and phi, Adam with clipping and scaling
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd.models import SVModel
    T, J, M, U = 40, 128, 4, 2
    ys = workloads()["sv_seed4_ys"]
    data = [ys[:T], ys[T:2 * T]]
    names = SVModel().param_names
    shared, unit = names[:2], names[2:]
    struct = panel_struct(SVModel(), data, shared, unit)
    th = osv.default_theta().astype(F32)
    sh0 = th[None, :2]
    un0 = np.stack([th[2:], th[2:]])[None]
    eta = pp.LearningRate({p: 0.02 for p in names})
    keys = tf.split(tf.key(78), M * U).reshape(1, M, U, 2)
    opt = pp.Adam(clip_norm=5.0, scale=True)
    nll, sh, un = F.panel_train(struct, sh0, un0, J, opt, M, eta, 0.97, keys, chunk_size=1)
    om = osv.SVModel()

    def vg(t, k, alpha, u):
        val, _, tt, *_ = mg.sv_mop_value_and_grad(t, data[u], k, J, alpha)
        return val, tt.grad.numpy().copy()

    est = om.to_est(th).astype(F32)
    onll, osh, oun = mg.panel_train(vg, est[:2], np.stack([est[2:], est[2:]]), np.arange(6), keys[0], T, M,
                                    np.full((M, 2), 0.02, F32), np.full((M, 4), 0.02, F32), 0.97,
                                    optimizer="adam", clip_norm=5.0, scale=True)
    full_o = np.concatenate([np.broadcast_to(osh[:, None, :], (M + 1, U, 2)), oun], axis=2)
    nat_o = om.from_est(full_o)
    np.testing.assert_allclose(nll[0, 1:], onll[1:], rtol=1e-4)
    np.testing.assert_allclose(sh[0], nat_o[:, 0, :2], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(un[0], nat_o[..., 2:], rtol=1e-4, atol=1e-5)


def test_panel_pomp_train_equals_the_functional_call():
    import pandas as pd
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd import random as prandom
    from pypomp_amd.models import LG
    th = lg1_theta()
    units = {f"u{s}": LG(A=np.array([[0.9]]), ys=lg_long(20, 50 + s)) for s in (0, 1, 2)}
    params = pp.PanelParameters([{
        "shared": pd.DataFrame({"shared": th[:2]}, index=["A11", "C11"]),
        "unit_specific": pd.DataFrame(np.repeat(th[2:, None], 3, axis=1), index=["Q11", "R11", "X0_1"],
                                      columns=list(units))}])
    P = pp.PanelPomp(units, params)
    eta = pp.LearningRate({p: 0.01 for p in LG_NAMES})
    P.train(J=64, M=3, eta=eta, key=prandom.key(3))
    new = tf.split(tf.key(3))[1]
    keys = tf.split(new, 3 * 3).reshape(1, 3, 3, 2)
    sh, un = params.arrays(list(units))
    nll, shh, unh = F.panel_train(P.to_struct(), sh, un, 64, pp.Adam(), 3, eta, 0.97, keys)
    r = P.results_history[-1]
    np.testing.assert_array_equal(r["shared_traces"][..., 0], -nll)
    np.testing.assert_array_equal(r["shared_traces"][..., 1:], shh)
    np.testing.assert_array_equal(r["unit_traces"][..., 1:], unh)
    assert np.isnan(r["unit_traces"][..., 0]).all()
    got_sh, got_un = P.theta.arrays(list(units))
    np.testing.assert_array_equal(got_sh, shh[:, -1])
    np.testing.assert_array_equal(got_un, unh[:, -1])
    tr = P.traces()
    assert set(tr["method"]) == {"train"} and len(tr) == 4 * 4


def test_unit_rows_equal_single_unit_structs():
    # mp_mop_grad and mp_mop with a unit per filter on a panel of uneven
    # lengths: each row equals the same filter on its unit alone
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    data = [lg_long(25, 60), lg_long(40, 61), lg_long(9, 62)]
    panel = panel_struct(LGModel(1, 1), data, [], LG_NAMES)
    unit = np.array([2, 0, 1, 1, 0], np.int32)
    rng = np.random.default_rng(9)
    th = (LGModel(1, 1).to_est(lg1_theta())[None] + F32(0.05) * rng.standard_normal((5, 5))).astype(F32)
    keys = tf.split(tf.key(64), 5)
    ll, g = F.mop_grad(panel, th, 80, 0.9, keys, K=7, unit=unit)
    v = F.mop(panel, th, 80, 0.9, keys, return_float64=True, unit=unit)
    for r, u in enumerate(unit):
        one = panel_struct(LGModel(1, 1), [data[u]], [], LG_NAMES)
        l1, g1 = F.mop_grad(one, th[r:r + 1], 80, 0.9, keys[r:r + 1], K=7)
        assert l1[0] == ll[r]
        np.testing.assert_array_equal(g1[0], g[r])
        assert F.mop(one, th[r:r + 1], 80, 0.9, keys[r:r + 1], return_float64=True)[0] == v[r]
    assert np.all(-v == ll)
    with pytest.raises(ValueError, match="needs the unit of every row"):
        F.mop_grad(panel, th, 80, 0.9, keys)
    with pytest.raises(ValueError, match="must lie in"):
        F.mop_grad(panel, th, 80, 0.9, keys, unit=np.full(5, 3))


@pytest.mark.parametrize("cs", [1, 3])
def test_panel_mop_matches_oracle(cs):
    # _chunked_panel_mop_internal and its value-and-gradient (core/algorithms/
    # mop.py:183-275): one launch over the units, the value bit-exact (same
    # per-unit float64 logliks, same float32 chunk sums), the gradient to the
    # adjoint's bar against autograd
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    U, T, J = 6, 30, 150
    data = [lg_long(T, 70 + u) for u in range(U)]
    shared, unit = ["C11", "A11"], ["X0_1", "R11", "Q11"]
    struct = panel_struct(LGModel(1, 1), data, shared, unit)
    omodel = olg.LGModel(olg.LGSpec(1, 1))
    names = shared + unit
    perm = np.array([names.index(p) for p in LG_NAMES])
    est = omodel.to_est(lg1_theta())[np.argsort(perm)]          # panel order
    rng = np.random.default_rng(11)
    sh = est[:2]
    un = (est[None, 2:] + F32(0.05) * rng.standard_normal((U, 3))).astype(F32)
    keys = tf.split(tf.key(90), U)
    v = F.panel_mop(struct, sh, un, J, 0.9, keys, chunk_size=cs)
    v2, gs, gu = F.panel_mop(struct, sh, un, J, 0.9, keys, chunk_size=cs, grad=True)

    def vg(t, k, u):
        val, _, g = mg.lg_mop_value_and_grad(omodel, t, data[u], k, J, 0.9)
        return val, g

    ov, ogs, ogu = mg.panel_mop(vg, sh, un, perm, keys, T, chunk_size=cs)
    assert v == ov and v2 == ov
    scale = max(np.abs(ogs).max(), np.abs(ogu).max())
    np.testing.assert_allclose(gs, ogs, rtol=1e-3, atol=1e-4 * scale)
    np.testing.assert_allclose(gu, ogu, rtol=1e-3, atol=1e-4 * scale)
