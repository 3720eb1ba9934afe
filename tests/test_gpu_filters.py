"""GPU parity of the filters: the reference's regression baselines, the oracle
restatement and the exact Kalman likelihood."""

import numpy as np
import pytest

from oracle import filters as fl
from oracle import kalman, lg as olg, threefry as tf
from tests.helpers import assert_close_tol, lg1_data, lg1_theta, workloads

pytestmark = pytest.mark.gpu
F32 = np.float32


def lg1_struct(seed=0):
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    return F.PompStruct(model=LGModel(1, 1), units=[Series.build(lg1_data(seed), np.arange(1, 5.0), 0.0, nstep=1)])


def lg2_struct(seed):
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    w = workloads()
    ys = w[f"lg2_seed{seed}_ys"]
    return (F.PompStruct(model=LGModel(2, 2, separate_init=True),
                         units=[Series.build(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)]),
            w[f"lg2_seed{seed}_theta"], ys)


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_pfilter_regression(golden, form):
    import pypomp_amd.functional as F
    keys = tf.split(tf.key(20260811), 1).reshape(1, 1, 2)
    out = F.pfilter(lg1_struct(), lg1_theta()[None], 12, keys, thresh=0.0, form=form)
    assert out["logLik"].dtype == np.float32
    assert_close_tol(out["logLik"], golden["test_pfilter_regression"]["logLik"])


def test_pfilter_diagnostics_regression(golden):
    import pypomp_amd.functional as F
    g = golden["test_pfilter_diagnostics_regression"]
    thetas = np.stack([lg1_theta(), (lg1_theta() * F32(1.05)).astype(F32)])
    keys = tf.split(tf.key(20260812), 2).reshape(2, 1, 2)
    out = F.pfilter(lg1_struct(), thetas, 12, keys, CLL=True, ESS=True, filter_mean=True, prediction_mean=True)
    assert_close_tol(out["logLik"].ravel(), g["logLik"])
    for name in ("CLL", "ESS", "filter_mean", "prediction_mean"):
        flat = out[name].reshape(2, -1)
        for r in range(2):
            for t in range(4):
                assert_close_tol(flat[r, t], g[f"{name}_rep{r}_t{t}"])


def test_mop_regression(golden):
    import pypomp_amd.functional as F
    v = F.mop(lg1_struct(), lg1_theta()[None], 12, 0.5, tf.split(tf.key(20260811), 1))
    assert_close_tol(v, golden["test_mop_regression"]["mop"])


def test_mif_regression(golden):
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    names = ["A11", "C11", "Q11", "R11", "X0_1"]
    rw = pp.RWSigma({n: 0.02 for n in names}).geometric_cooling(0.5)
    th = np.repeat(lg1_theta()[None, None, :], 12, axis=1)
    lls, tr, _ = F.mif(lg1_struct(), th, rw, 3, 12, tf.split(tf.key(20260811), 1))
    g = golden["test_mif_regression"]
    assert_close_tol(lls[:, -1], g["final_loglik"])
    assert_close_tol(tr[:, -1, :], g["final_theta"])


def test_mif_full_trace_regression(golden):
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    names = ["A11", "C11", "Q11", "R11", "X0_1"]
    g = golden["test_mif_full_trace_regression"]
    rw = pp.RWSigma({n: 0.02 for n in names}).geometric_cooling(0.5)
    thetas = np.stack([lg1_theta(), (lg1_theta() * F32(1.05)).astype(F32)])
    th = np.repeat(thetas[:, None, :], 12, axis=1)
    lls, tr, _ = F.mif(lg1_struct(), th, rw, 4, 12, tf.split(tf.key(20260812), 2), n_monitors=1)
    for r in range(2):
        for m in range(4):
            assert_close_tol(lls[r, m], g[f"loglik_rep{r}_m{m}"])
        for m in range(5):
            for p, name in enumerate(names):
                assert_close_tol(tr[r, m, p], g[f"{name}_rep{r}_m{m}"])


def test_panel_pfilter_units_regression(golden):
    import pypomp_amd.functional as F
    keys = tf.split(tf.key(20260811), 2)
    got = [F.pfilter(lg1_struct(s), lg1_theta()[None], 12, keys[u].reshape(1, 1, 2))["logLik"][0, 0]
           for u, s in enumerate((0, 1))]
    assert_close_tol(got, golden["test_panel_pfilter_regression"]["logLik"])


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_lg2_matches_oracle(form):
    import pypomp_amd.functional as F
    struct, theta, ys = lg2_struct(1)
    keys = tf.split(tf.key(1234), 3)
    flags = dict(CLL=True, ESS=True, filter_mean=True, prediction_mean=True)
    out = F.pfilter(struct, theta[None], 1000, keys.reshape(1, 3, 2), form=form, **flags)
    model = olg.LGModel(olg.LGSpec(2, 2, separate_init=True))
    ser = fl.make_series(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)
    for r in range(3):
        o = fl.pfilter(model, theta, ser, keys[r], 1000, diagnostics=True)
        # identical trajectories and an order-independent exp-sum: the float64
        # log-likelihood is bit-exact
        assert float(out["logLik64"][0, r]) == o["logLik64"], (r, out["logLik64"][0, r], o["logLik64"])
        np.testing.assert_array_equal(out["CLL"][0, r], np.array(o["CLL"], dtype=np.float32))
        # identical particle trajectories: the same predicted and filtered means
        # and ESS (float64 sums on the GPU, float32 sums in the oracle)
        np.testing.assert_allclose(out["prediction_mean"][0, r], np.array(o["prediction_mean"]),
                                   rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(out["filter_mean"][0, r], np.array(o["filter_mean"]), rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(out["ESS"][0, r], np.array(o["ESS"]), rtol=1e-5)


def test_plain_and_batched_choose_the_same_ancestors():
    import pypomp_amd.functional as F
    struct, theta, _ = lg2_struct(1)
    keys = tf.split(tf.key(99), 4).reshape(1, 4, 2)
    a = F.pfilter(struct, theta[None], 5000, keys, CLL=True, form="batched")
    b = F.pfilter(struct, theta[None], 5000, keys, CLL=True, form="plain")
    # identical resampling decisions and exact fixed-point totals in both forms
    np.testing.assert_array_equal(a["logLik64"], b["logLik64"])
    np.testing.assert_array_equal(a["CLL"], b["CLL"])


def test_lg2_against_kalman():
    import pypomp_amd.functional as F
    struct, theta, ys = lg2_struct(1)
    A = np.array([[np.cos(0.2), -np.sin(0.2)], [np.sin(0.2), np.cos(0.2)]])
    spec = olg.LGSpec(2, 2, separate_init=True)
    _, C, LQ, LR, X0, L0 = spec.unpack(theta)
    exact = kalman.kalman_loglik(ys, A, C, LQ.astype(float) @ LQ.T.astype(float),
                                 LR.astype(float) @ LR.T.astype(float), X0.astype(float),
                                 L0.astype(float) @ L0.T.astype(float))
    reps = 64
    out = F.pfilter(struct, theta[None], 1000, tf.split(tf.key(77), reps).reshape(1, reps, 2))
    ll = out["logLik64"].ravel()
    se = ll.std(ddof=1) / np.sqrt(reps)
    # the particle estimate of the likelihood is unbiased; its log has a small negative bias
    assert abs(ll.mean() - exact) < 4 * se + 0.05, (ll.mean(), exact, se)


def test_thresh_skips_resampling_like_oracle():
    import pypomp_amd.functional as F
    struct, theta, ys = lg2_struct(1)
    keys = tf.split(tf.key(5), 2)
    out = F.pfilter(struct, theta[None], 400, keys.reshape(1, 2, 2), thresh=0.5, ESS=True)
    model = olg.LGModel(olg.LGSpec(2, 2, separate_init=True))
    ser = fl.make_series(ys, np.arange(1, ys.shape[0] + 1.0), 0.0, nstep=1)
    for r in range(2):
        o = fl.pfilter(model, theta, ser, keys[r], 400, thresh=0.5, diagnostics=True)
        assert float(out["logLik64"][0, r]) == o["logLik64"]
        np.testing.assert_allclose(out["ESS"][0, r], np.array(o["ESS"]), rtol=1e-3)


def test_j_equals_one_and_workspace_errors():
    import pypomp_amd.functional as F
    from pypomp_amd import _native
    struct, theta, _ = lg2_struct(1)
    out = F.pfilter(struct, theta[None], 1, tf.split(tf.key(3), 1).reshape(1, 1, 2))
    assert np.isfinite(out["logLik64"]).all()
    with pytest.raises(ValueError):
        F.pfilter(struct, theta[None], 0, tf.split(tf.key(3), 1).reshape(1, 1, 2))


def test_launch_shapes_do_not_change_results():
    # every tunable launch shape computes the same bits
    import pypomp_amd.functional as F
    from pypomp_amd import autotune
    struct, theta, ys = lg2_struct(1)
    keys = tf.split(tf.key(77), 3).reshape(1, 3, 2)
    saved = {n: autotune.get(n) for n in autotune.PARAMS}
    try:
        outs = []
        for nt in (256, 512, 768):
            autotune.set("batched_block", nt)
            outs.append(F.pfilter(struct, theta[None], 5000, keys, form="batched"))
        for sr, pr, nt in ((2, 1, 256), (8, 4, 256), (16, 2, 256), (4, 2, 512), (16, 1, 512)):
            autotune.set("plain_scan_rows", sr)
            autotune.set("plain_prop_rows", pr)
            autotune.set("plain_scan_threads", nt)
            autotune.set("plain_prop_threads", nt)
            outs.append(F.pfilter(struct, theta[None], 5000, keys, form="plain"))
        for o in outs[1:]:
            np.testing.assert_array_equal(o["logLik64"], outs[0]["logLik64"])
        best = autotune.tune_pfilter(struct, theta, 5000, 3, form="batched")
        assert best["batched_block"] in (256, 512, 768) and best["ms"] > 0
    finally:
        for n, v in saved.items():
            autotune.set(n, v)


def test_plain_xcd_chunk_map_does_not_change_results():
    # Propagate blocks taking the chunks of the scan tiles on their own XCD (the
    # default) or chunks in block order compute the same bits, with 1, 2, 4 and
    # 16 propagate chunks per scan tile; all equal the batched form.
    import pypomp_amd.functional as F
    from pypomp_amd import _native, autotune
    lib = _native.lib()
    struct, theta, _ = lg2_struct(1)
    keys = tf.split(tf.key(93), 2).reshape(1, 2, 2)
    J = 200_000
    ref = F.pfilter(struct, theta[None], J, keys, form="batched")
    names = ("plain_scan_rows", "plain_scan_threads", "plain_prop_rows", "plain_prop_threads")
    saved = {n: autotune.get(n) for n in names}
    try:
        for shape in ((1, 256, 1, 256), (2, 1024, 2, 512), (4, 1024, 2, 512), (4, 1024, 1, 256)):
            for n, v in zip(names, shape):
                autotune.set(n, v)
            for xmap in (1, 0):
                _native.check(lib.mp_tune_set(_native.MP_TUNE_PLAIN_XCD_MAP, xmap))
                out = F.pfilter(struct, theta[None], J, keys, form="plain")
                np.testing.assert_array_equal(out["logLik64"], ref["logLik64"], err_msg=f"{shape} map {xmap}")
    finally:
        for n, v in saved.items():
            autotune.set(n, v)
        _native.check(lib.mp_tune_set(_native.MP_TUNE_PLAIN_XCD_MAP, 1))


@pytest.mark.parametrize("J", [32767, 32768, 32769])
def test_plain_xcd_chunk_map_edges(J):
    # The XCD-local order needs at least 8 q = 32 full propagate chunks of 1024
    # particles: J = 32767 (32 chunks, the last one short), 32768 (exactly 32
    # full chunks) and 32769 (33 chunks: 64 blocks, 31 of them empty).
    import pypomp_amd.functional as F
    struct, theta, _ = lg2_struct(1)
    keys = tf.split(tf.key(95), 2).reshape(1, 2, 2)
    plain = F.pfilter(struct, theta[None], J, keys, form="plain", CLL=True)
    batched = F.pfilter(struct, theta[None], J, keys, form="batched", CLL=True)
    np.testing.assert_array_equal(plain["logLik64"], batched["logLik64"])
    np.testing.assert_array_equal(plain["CLL"], batched["CLL"])


def _with_plain_tuning(settings, fn):
    from pypomp_amd import _native, autotune
    lib = _native.lib()
    names = ("plain_scan_rows", "plain_scan_threads", "plain_lanes")
    saved = {n: autotune.get(n) for n in names}
    lds = lib.mp_tune_get(_native.MP_TUNE_PLAIN_LDS_TILES)
    try:
        for n, v in settings.items():
            if n == "lds_tiles":
                _native.check(lib.mp_tune_set(_native.MP_TUNE_PLAIN_LDS_TILES, v))
            else:
                autotune.set(n, v)
        return fn()
    finally:
        for n, v in saved.items():
            autotune.set(n, v)
        _native.check(lib.mp_tune_set(_native.MP_TUNE_PLAIN_LDS_TILES, lds))


@pytest.mark.parametrize("settings", [
    {"lds_tiles": 0},                                                   # tile ends by a separate launch
    {"plain_scan_rows": 1, "plain_scan_threads": 256},                  # 782 tiles of 256 in LDS
    {"plain_scan_rows": 1, "plain_scan_threads": 256, "lds_tiles": 0},  # 782 tiles, global ends
    {"plain_scan_rows": 1, "plain_scan_threads": 256, "lds_tiles": 100},  # rows widened to 8: 98 tiles
])
def test_plain_tile_ends_do_not_change_results(settings):
    # The propagate blocks take the prefix of the scan tiles before each
    # position from the tiles' totals, formed in LDS by every block (the
    # default) or written to global memory by a separate launch per step (more
    # tiles than the LDS limit); the scan threads widen their runs to stay under
    # the limit.  Every variant gives the bits of the batched form.
    import pypomp_amd.functional as F
    struct, theta, _ = lg2_struct(1)
    keys = tf.split(tf.key(91), 2).reshape(1, 2, 2)
    J = 200_000
    ref = F.pfilter(struct, theta[None], J, keys, form="batched", CLL=True)
    out = _with_plain_tuning(settings, lambda: F.pfilter(struct, theta[None], J, keys, form="plain", CLL=True))
    np.testing.assert_array_equal(out["logLik64"], ref["logLik64"])
    np.testing.assert_array_equal(out["CLL"], ref["CLL"])


@pytest.mark.parametrize("nsets,reps,thresh", [(1, 3, 0.0), (2, 2, 0.0), (1, 5, 0.0), (1, 5, 0.5)])
def test_plain_lanes_do_not_change_results(nsets, reps, thresh):
    # With lanes the plain form runs replicate filters in groups, each on its
    # own stream with its own buffers (a short last group leaves lanes idle):
    # every output, diagnostics included, keeps the bits of one lane.
    import pypomp_amd.functional as F
    struct, theta, _ = lg2_struct(1)
    thetas = np.stack([theta, theta * F32(1.05)])[:nsets]
    keys = tf.split(tf.key(93), nsets * reps).reshape(nsets, reps, 2)
    J = 150_000

    def run(lanes):
        return _with_plain_tuning({"plain_lanes": lanes}, lambda: F.pfilter(
            struct, thetas, J, keys, thresh=thresh, form="plain", CLL=True, ESS=True, filter_mean=True,
            prediction_mean=True))

    one, two = run(1), run(2)
    for k in ("logLik64", "CLL", "ESS", "filter_mean", "prediction_mean"):
        np.testing.assert_array_equal(two[k], one[k], err_msg=k)
    four = run(4)  # three, four or five filters: lanes idle, all busy, or a last group of one
    for k in ("logLik64", "CLL", "ESS", "filter_mean", "prediction_mean"):
        np.testing.assert_array_equal(four[k], one[k], err_msg=k)
    ref = F.pfilter(struct, thetas, J, keys, thresh=thresh, form="batched", CLL=True)
    np.testing.assert_array_equal(two["logLik64"], ref["logLik64"])
    np.testing.assert_array_equal(two["CLL"], ref["CLL"])


def test_plain_lanes_replay_from_a_hip_graph():
    # a plain-form call with four lanes, captured on a side stream (the lanes
    # fork from it and join back by events), replays with the eager bits; eager
    # calls run between the replays
    import torch
    from pypomp_amd import _native
    from pypomp_amd.functional import _keys_tensor
    struct, theta, _ = lg2_struct(1)
    lib = _native.lib()
    dev = _native.require_device()
    R, J = 4, 50_000
    ser, md = struct.device_series(), struct.model.desc()
    keys = _keys_tensor(tf.split(tf.key(99), R), dev)
    th = torch.from_numpy(np.repeat(theta[None], R, 0)).to(dev)
    ll = torch.empty(R, dtype=torch.float64, device=dev)
    n = int(lib.mp_filter_workspace_size(_native.MP_ALGO_PF, md, ser.desc, R, J, _native.MP_FORM_PLAIN))
    ws = torch.empty(n, dtype=torch.uint8, device=dev)

    def call():
        _native.check(lib.mp_pfilter(md, ser.desc, _native.ptr(th), struct.model.d, _native.ptr(keys), None, R, J,
                                     0.0, 0, _native.MP_FORM_PLAIN, _native.ptr(ll), None, None, None, None,
                                     _native.ptr(ws), n, _native.stream_ptr()))

    assert lib.mp_tune_get(_native.MP_TUNE_PLAIN_LANES) == 4
    call()
    torch.cuda.synchronize()
    eager = ll.clone()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    with torch.cuda.graph(g, stream=side):
        call()
    for _ in range(3):
        ll.fill_(0.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(ll, eager)
        call()
        torch.cuda.synchronize()
        assert torch.equal(ll, eager)


@pytest.mark.parametrize("J", [1, 7, 1000, 4097])
def test_plain_four_lanes_small_filters_equal_batched(J):
    # four filters in four lanes take the smaller workgroups (scan 256 x 16,
    # propagate 256 x 4) at any J, down to one particle
    import pypomp_amd.functional as F
    struct, theta, _ = lg2_struct(1)
    keys = tf.split(tf.key(97 + J), 4).reshape(1, 4, 2)
    plain = F.pfilter(struct, theta[None], J, keys, form="plain", CLL=True)
    batched = F.pfilter(struct, theta[None], J, keys, form="batched", CLL=True)
    np.testing.assert_array_equal(plain["logLik64"], batched["logLik64"])
    np.testing.assert_array_equal(plain["CLL"], batched["CLL"])


@pytest.mark.parametrize("lds_tiles", [1024, 0])
def test_plain_wide_ancestor_windows_match_batched_and_oracle(lds_tiles):
    # An outlying observation leaves a few hundred particles with non-zero
    # fixed-point weights, scattered over J = 100000 indices: a block's comb
    # then spans thousands of particles, more than its LDS window holds, and
    # the per-particle search over global memory runs.  Plain (both tile-end
    # paths), batched and the oracle pick the same ancestors.
    import pypomp_amd.functional as F
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    _, theta, ys = lg2_struct(1)
    ys = ys[:6].copy()
    ys[2] = ys[2] + F32(4.0)
    times = np.arange(1, ys.shape[0] + 1.0)
    struct = F.PompStruct(model=LGModel(2, 2, separate_init=True), units=[Series.build(ys, times, 0.0, nstep=1)])
    keys = tf.split(tf.key(97), 1).reshape(1, 1, 2)
    J = 100_000
    batched = F.pfilter(struct, theta[None], J, keys, form="batched", CLL=True)
    plain = _with_plain_tuning({"lds_tiles": lds_tiles},
                               lambda: F.pfilter(struct, theta[None], J, keys, form="plain", CLL=True))
    np.testing.assert_array_equal(plain["logLik64"], batched["logLik64"])
    np.testing.assert_array_equal(plain["CLL"], batched["CLL"])
    exp = fl.pfilter(olg.LGModel(olg.LGSpec(2, 2, separate_init=True)), theta,
                     fl.make_series(ys, times, 0.0, nstep=1), keys[0, 0], J, diagnostics=True)
    assert float(plain["logLik64"][0, 0]) == exp["logLik64"]
    np.testing.assert_array_equal(plain["CLL"][0, 0], np.array(exp["CLL"], dtype=F32))


def test_pomp_pickle_continues_the_key_stream_and_merge_joins_runs():
    # core/pomp.py:542-826: an unpickled Pomp equals the original and draws the
    # same next keys; merge joins the parameter sets and results of two runs
    import pickle
    import pypomp_amd as pp
    from pypomp_amd.models.lg import LG
    names = ["A11", "C11", "Q11", "R11", "X0_1"]
    rw = pp.RWSigma({n: 0.02 for n in names}).geometric_cooling(0.5)
    P = LG(A=np.array([[0.9]]), ys=lg1_data(0))
    P.fresh_key = pp.random.key(5)
    P.pfilter(J=50, reps=2)
    Q = pickle.loads(pickle.dumps(P))
    assert P == Q
    P.pfilter(J=50, reps=2)
    Q.pfilter(J=50, reps=2)
    np.testing.assert_array_equal(P.results_history[-1]["logLiks"], Q.results_history[-1]["logLiks"])
    assert P == Q
    P.mif(J=20, M=2, rw_sd=rw)
    Q.mif(J=20, M=2, rw_sd=rw)
    R = LG(A=np.array([[0.5]]), ys=lg1_data(0))
    R.fresh_key = pp.random.key(6)
    for _ in range(2):
        R.pfilter(J=50, reps=2)
    R.mif(J=20, M=2, rw_sd=rw)
    M = type(P).merge(P, R)
    assert len(M.theta) == 2 and len(M.results_history) == 3
    res = M.results(0)
    assert list(res["theta_idx"]) == [0, 1]
    np.testing.assert_array_equal(res["logLik"].to_numpy()[:1], P.results(0)["logLik"].to_numpy())
    np.testing.assert_array_equal(res["logLik"].to_numpy()[1:], R.results(0)["logLik"].to_numpy())
    tr = M.traces()
    assert set(tr["theta_idx"]) == {0, 1} and set(tr["method"]) == {"pfilter", "mif"}
    with pytest.raises(ValueError, match="same number of entries"):
        type(P).merge(P, LG(A=np.array([[0.5]]), ys=lg1_data(0)))


def _lg_oo():
    import pypomp_amd as pp
    from pypomp_amd.models.lg import LG
    return LG(A=np.array([[0.9]]), T=5, key=pp.random.key(0))


def _reversed(theta):
    return [{k: row[k] for k in reversed(list(row))} for row in theta]


def _oo_pfilter(P, theta):
    P.pfilter(J=64, key=tf.key(11), theta=theta, CLL=True)
    return np.asarray(P.results_history[-1]["logLiks"])


def _oo_mif(P, theta):
    import pypomp_amd as pp
    rw = pp.RWSigma({n: 0.02 for n in P.canonical_param_names}).geometric_cooling(0.5)
    P.mif(J=64, M=2, rw_sd=rw, key=tf.key(11), theta=theta, n_monitors=1)
    return np.asarray(P.results_history[-1]["traces"])


def _oo_train(P, theta):
    import pypomp_amd as pp
    eta = pp.LearningRate({n: 0.2 for n in P.canonical_param_names})
    P.train(J=64, M=2, eta=eta, key=tf.key(11), theta=theta)
    return np.asarray(P.results_history[-1]["traces"])


def _oo_simulate(P, theta):
    X, Y = P.simulate(key=tf.key(11), theta=theta, nsim=1)
    return np.concatenate([X.to_numpy().ravel(), Y.to_numpy().ravel()])


@pytest.mark.parametrize("run", [_oo_pfilter, _oo_mif, _oo_train, _oo_simulate],
                         ids=["pfilter", "mif", "train", "simulate"])
def test_param_order_invariance(run):
    # tests/test_properties/test_invariants.py:145-171: results do not depend on
    # the order theta's keys were supplied in (train with Adam: the reference's
    # Newton needs the MOP Hessian)
    P = _lg_oo()
    theta = [dict(t) for t in P.theta]
    base = run(P, theta)
    P.results_history.clear()
    perm = run(P, _reversed(theta))
    np.testing.assert_array_equal(base, perm)


def test_oo_cll_sum_and_mif_trace_start():
    # tests/test_properties/test_invariants.py:50-75: the conditional
    # log-likelihoods sum to the log-likelihood, and iteration 0 of the IF2
    # trace is the parameter vector IF2 started from
    import pypomp_amd as pp
    P = _lg_oo()
    P.pfilter(J=64, key=tf.key(12), CLL=True, reps=2)
    r = P.results_history[-1]
    np.testing.assert_allclose(np.asarray(r["CLL"], np.float64).sum(axis=-1), r["logLiks"], rtol=1e-5, atol=1e-5)
    names = P.canonical_param_names
    start = np.array([[t[n] for n in names] for t in P.theta], np.float32)
    rw = pp.RWSigma({n: 0.02 for n in names}).geometric_cooling(0.5)
    P.mif(J=64, M=2, rw_sd=rw, key=tf.key(12), n_monitors=1)
    tr = P.results_history[-1]["traces"]
    np.testing.assert_allclose(tr[:, 0, 1:], start, rtol=1e-6, atol=1e-6)


def test_object_api_equals_the_functional_calls():
    # tests/test_properties/test_layer_parity.py:51-135: Pomp.pfilter, Pomp.mif
    # and Pomp.simulate equal hand-built functional calls with the key derived
    # from the given key (split(key)[1]), bit for bit
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    from pypomp_amd import random as prandom
    P = _lg_oo()
    names = P.canonical_param_names
    th = P.model.theta_array(P.theta)
    key = tf.key(13)
    new = prandom.split(key)[1]

    P.pfilter(J=64, key=key, reps=2, CLL=True, ESS=True)
    r = P.results_history[-1]
    exp = F.pfilter(P.to_struct(), th, 64, prandom.split(new, 2).reshape(1, 2, 2), 0.0, CLL=True, ESS=True)
    np.testing.assert_array_equal(r["logLiks"], exp["logLik"])
    np.testing.assert_array_equal(r["CLL"], exp["CLL"])
    np.testing.assert_array_equal(r["ESS"], exp["ESS"])

    rw = pp.RWSigma({n: 0.02 for n in names}).geometric_cooling(0.5)
    th = P.model.theta_array(P.theta)
    P.mif(J=64, M=2, rw_sd=rw, key=key)
    tr = P.results_history[-1]["traces"]
    lls, traces, _ = F.mif(P.to_struct(), np.repeat(th[:, None, :], 64, axis=1), rw, 2, 64,
                           prandom.split(new, 1), 0.0)
    assert np.all(np.isnan(tr[:, 0, 0]))
    np.testing.assert_array_equal(tr[:, 1:, 0], lls)
    np.testing.assert_array_equal(tr[:, :, 1:], traces)

    th = P.model.theta_array(P.theta)
    X, Y = P.simulate(key=key, nsim=3)
    Xs, Ys = F.simulate(P.to_struct(), th, 3, prandom.split(new, 1))
    np.testing.assert_array_equal(Y.drop(columns=["theta_idx", "sim", "time"]).to_numpy(),
                                  np.asarray(Ys).reshape(-1, Ys.shape[-1]))
    np.testing.assert_array_equal(X.drop(columns=["theta_idx", "sim", "time"]).to_numpy(),
                                  np.asarray(Xs).reshape(-1, Xs.shape[-1]))


@pytest.mark.parametrize("schedule", ["cosine", "hyperbolic", "custom"])
@pytest.mark.parametrize("graph", [False, True])
def test_mif_cooling_schedules_match_oracle(schedule, graph):
    # the non-geometric schedules (core/rw_sigma.py:193-225, 308-376) reach the
    # kernel as per-observation factor tables: traces and final swarm equal the
    # oracle's mif, eagerly and replayed from a HIP graph
    import pypomp_amd as pp
    import pypomp_amd.functional as F
    names = ["A11", "C11", "Q11", "R11", "X0_1"]
    base = pp.RWSigma({n: 0.05 for n in names}, init_names=["X0_1"])
    fn = lambda nt, m, ntimes: 0.8 ** (m + nt / ntimes)  # noqa: E731
    rw = {"cosine": base.cosine_cooling(0.2, 4), "hyperbolic": base.hyperbolic_cooling(0.7),
          "custom": base.custom_cooling(fn)}[schedule]
    init = np.array([n == "X0_1" for n in names], np.float64)
    kw = {"cosine": dict(c=0.2, M=4), "hyperbolic": dict(s=0.7), "custom": dict(fn=fn)}[schedule]
    orw = fl.RWSigma(np.full(5, 0.05), init, a=0.5, cooling=schedule, **kw)
    J, M = 40, 4
    th = np.repeat(lg1_theta()[None, None, :], J, axis=1)
    keys = tf.split(tf.key(44), 2)
    lls, tr, final = F.mif(lg1_struct(), np.repeat(th, 2, axis=0), rw, M, J, keys, graph=graph)
    model = olg.LGModel(olg.LGSpec(1, 1))
    ser = fl.make_series(lg1_data(0), np.arange(1, 5.0), 0.0, nstep=1)
    for r in range(2):
        ell, etr, efinal = fl.mif(model, th[0], ser, keys[r], J, M, orw)
        # the swarm bit for bit; its means (float64 sums here, float32 in the
        # oracle) to the parity bar
        np.testing.assert_array_equal(final[r], efinal)
        assert_close_tol(tr[r], etr)
        np.testing.assert_allclose(lls[r], ell, rtol=1e-5)
    # another schedule gives another path
    _, tr_geo, _ = F.mif(lg1_struct(), np.repeat(th, 2, axis=0), base.geometric_cooling(0.5), M, J, keys)
    assert not np.array_equal(tr, tr_geo)


def test_auto_form_runs_diagnostics_in_the_plain_form():
    # Pomp.pfilter(J = 2^18, ESS, means) takes the plain kernels ("auto": a few
    # filters of many particles) and equals the batched form
    import pypomp_amd.functional as F
    from pypomp_amd import _native
    from pypomp_amd.models import LGModel
    from pypomp_amd.series import Series
    _, theta, ys = lg2_struct(2)
    struct = F.PompStruct(model=LGModel(2, 2, separate_init=True),
                          units=[Series.build(ys[:40], np.arange(1, 41.0), 0.0, nstep=1)])
    J = 1 << 18
    assert F.choose_form(2, J) == "plain" and F.choose_form(2, 4096) == "batched"
    assert F.choose_form(512, J) == "batched" and F.choose_form(2, J, multi_unit=True) == "batched"
    keys = tf.split(tf.key(61), 2).reshape(1, 2, 2)
    flags = dict(CLL=True, ESS=True, filter_mean=True, prediction_mean=True)
    _native.profile_enable(True)
    _native.profile_read()
    a = F.pfilter(struct, theta[None], J, keys, **flags)
    prof = _native.profile_read()
    _native.profile_enable(False)
    assert prof["plain_prop"][1] == 2 * 40 and prof["batched"][1] == 0
    b = F.pfilter(struct, theta[None], J, keys, form="batched", **flags)
    np.testing.assert_array_equal(a["logLik64"], b["logLik64"])
    np.testing.assert_array_equal(a["CLL"], b["CLL"])
    for k in ("ESS", "filter_mean", "prediction_mean"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-6, atol=1e-7)
    assert np.all(a["ESS"] > 0) and np.all(a["ESS"] <= J * 1.0001)
