"""Level planning on the device against the host planner: every tiled precision form and the
growth parameters that change a level's plan (leaf cap, depth, monotone bounds, minimum leaf
size, weights, missing values, random thresholds, softmax) must give the same model bytes
and the same fit statistics, and failures must raise the same errors. MPSB_HOST_PLAN=1 makes
a fit plan on the host; MPSB_TREE_GRAPH=1 replays a captured graph of a device-planned tree's
launches instead of launching every tree directly."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")


def _with_env(name, value, fn):
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        return fn()
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


def _with_host_plan(host, fn):
    return _with_env("MPSB_HOST_PLAN", "1" if host else "0", fn)


def _data(rows=40_000, feats=12, seed=0, missing=True):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    if missing:
        X[rng.random((rows, feats)) < 0.05] = np.nan
        X[:, 3] = np.round(X[:, 3] * 2) + 0.0  # a few distinct values
    y = np.sin(np.nan_to_num(X[:, 0])) + np.nan_to_num(X[:, 1]) * np.nan_to_num(X[:, 2])
    y += 0.1 * rng.standard_normal(rows)
    return X, y


def _both(X, y, w=None, classes=0, max_bins=64, **kw):
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, max_bins)
    p = N._TrainingParameters(kw.pop("n_estimators", 4), kw.pop("learning_rate", 0.3), max_bins,
                              kw.pop("max_depth", 6), kw.pop("min_samples_leaf", 5), 0.0, 1.0, **kw)
    fit = lambda: N._train_binned(s, ds, y, w, p, classes)  # noqa: E731
    host = _with_host_plan(True, fit)
    dev = _with_host_plan(False, fit)
    return host, dev


def _same(host, dev):
    assert dev.serialize() == host.serialize()
    a, b = host.fit_stats(), dev.fit_stats()
    for k in ("histogram_rows", "histogram_bytes", "histogram_launches", "trees"):
        assert a[k] == b[k], k


@pytest.mark.parametrize("precision", ["int16", "int8", "fp64_fixed"])
def test_forms_match_the_host_plan(precision):
    X, y = _data()
    _same(*_both(X, y, precision=precision, rounding_seed=5))


def test_fp32_form_matches_the_host_plan_within_rounding():
    # FP32 LDS atomics add in a run-dependent order, so two fits agree to rounding only.
    X, y = _data()
    host, dev = _both(X, y, precision="fp32")
    assert dev.tree_count == host.tree_count
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, 64)
    a = N._RegressionModel(host, ds, s).predict(X)
    b = N._RegressionModel(dev, ds, s).predict(X)
    assert np.max(np.abs(a - b)) < 1e-3 * np.std(y)


@pytest.mark.parametrize("kw", [
    {"max_leaves": 11},
    {"max_leaves": 2, "max_depth": 4},
    {"max_depth": 1},
    {"max_depth": 10, "min_samples_leaf": 1},
    {"min_samples_leaf": 400},
    {"monotonic_constraints": [1, -1] + [0] * 10},
    {"reg_alpha": 0.5, "max_delta_step": 0.05, "gamma": 0.01},
    {"split_strategy": "random_threshold", "random_seed": 11},
    {"objective": "binary_logistic"},
    {"objective": "poisson"},
    {"objective": "quantile", "objective_alpha": 0.3},
])
def test_growth_parameters_match_the_host_plan(kw):
    X, y = _data(rows=20_000, seed=1)
    if kw.get("objective") == "binary_logistic":
        y = (y > 0).astype(np.float64)
    if kw.get("objective") == "poisson":
        y = np.exp(np.clip(y, -3, 3))
    _same(*_both(X, y, precision="int16", **kw))


def test_weights_and_no_missing_match_the_host_plan():
    X, y = _data(rows=30_000, seed=2, missing=False)
    w = np.random.default_rng(3).uniform(0.2, 3.0, size=y.size)
    _same(*_both(X, y, w, precision="fp64_fixed"))


def test_softmax_matches_the_host_plan():
    X, y = _data(rows=25_000, seed=4)
    cls = np.digitize(y, [-0.5, 0.5]).astype(np.float64)
    _same(*_both(X, cls, classes=3, precision="int16", n_estimators=3))


def test_tiny_data_and_single_rows_per_leaf():
    X, y = _data(rows=300, feats=5, seed=6)
    _same(*_both(X, y, precision="int8", max_depth=9, min_samples_leaf=1, n_estimators=6))


def test_overflow_error_is_the_same():
    X, _ = _data(rows=5_000, seed=7)
    y = np.full(X.shape[0], 1e308)
    y[::2] = -1e308
    errs = []
    for host in (True, False):
        with pytest.raises(N.TrainingError) as e:
            _with_host_plan(host, lambda: _both_one(X, y))
        errs.append(str(e.value))
    assert errs[0] == errs[1]


def _both_one(X, y):
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, 32)
    p = N._TrainingParameters(3, 0.3, 32, 4, 5, 0.0, 1.0, precision="int16")
    return N._train_binned(s, ds, y, None, p)


@pytest.mark.parametrize("classes", [0, 3])
def test_tree_graph_replays_match_direct_launches(classes):
    # Untimed trees replay one captured graph (several replays, with the gradient passes and
    # the timed trees' direct launches between them); every tree must equal the direct launch.
    X, y = _data(rows=30_000, seed=8)
    if classes:
        y = np.digitize(y, [-0.5, 0.5]).astype(np.float64)
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, 64)
    p = N._TrainingParameters(12 if classes == 0 else 5, 0.3, 64, 6, 5, 0.0, 1.0, precision="int16",
                              rounding_seed=3)
    fit = lambda: N._train_binned(s, ds, y, None, p, classes)  # noqa: E731
    direct = _with_env("MPSB_TREE_GRAPH", "0", fit)
    replay = _with_env("MPSB_TREE_GRAPH", "1", fit)
    twice = _with_env("MPSB_TREE_GRAPH", "1", fit)
    assert replay.serialize() == direct.serialize() == twice.serialize()
    a, b = direct.fit_stats(), replay.fit_stats()
    for k in ("histogram_rows", "histogram_bytes", "histogram_launches", "trees", "levels"):
        if k in a:
            assert a[k] == b[k], k


@pytest.mark.parametrize("precision", ["int16", "fp64"])
def test_wide_margin_loads_match_buffer_loads(precision):
    # Softmax margins beyond 4 GiB are read with 64-bit addresses instead of one buffer
    # resource; MPSB_WIDE_MARGINS=1 forces that path at a small size. Both must give the same
    # model bytes (the same margins, exponentials and class order), with weights, through the
    # two-pass (int16) and single-pass (fp64) gradient kernels.
    X, y = _data(rows=20_000, seed=9)
    cls = np.digitize(y, [-1.0, -0.3, 0.3, 1.0]).astype(np.float64)
    w = np.random.default_rng(4).uniform(0.5, 2.0, size=y.size)
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, 64)
    p = N._TrainingParameters(3, 0.3, 64, 5, 5, 0.0, 1.0, precision=precision, rounding_seed=2)
    fit = lambda: N._train_binned(s, ds, cls, w, p, 5)  # noqa: E731
    wide = _with_env("MPSB_WIDE_MARGINS", "1", fit)
    narrow = _with_env("MPSB_WIDE_MARGINS", "0", fit)
    assert wide.tree_count == 15
    assert wide.serialize() == narrow.serialize()


@pytest.mark.parametrize("precision", ["int16", "int8"])
@pytest.mark.parametrize("feats", [100, 130])
def test_gradient_only_root_matches_the_packed_root(precision, feats):
    # Unit-Hessian integer fits build every root after the first from gradients alone (4-byte
    # cells, 64-feature tiles) and take its counts from the first tree: the model bytes must
    # equal the packed root's (MPSB_ROOT_GONLY=0), over tiles of 16, 12 and 4 row words, narrow
    # and wide features and missing values.
    rng = np.random.default_rng(feats)
    rows = 30_000
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    X[:, : feats // 5] = rng.integers(0, 7, (rows, feats // 5)).astype(np.float32)
    X[rng.random(X.shape) < 0.1] = np.nan
    y = np.sin(np.nan_to_num(X[:, -1])) + np.nan_to_num(X[:, 0]) * np.nan_to_num(X[:, -2])
    y += 0.1 * rng.standard_normal(rows)
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, 255)  # 255 bins + the missing code: one-byte codes
    assert ds.code_bytes == 1
    p = N._TrainingParameters(6, 0.3, 255, 5, 5, 0.0, 1.0, precision=precision, rounding_seed=11)
    fit = lambda: N._train_binned(s, ds, y, None, p)  # noqa: E731
    packed = _with_env("MPSB_ROOT_GONLY", "0", lambda: _with_host_plan(False, fit))
    assert packed.fit_stats()["root_gradient_only_builds"] == 0
    for host in (False, True):  # both planners take gradient-only roots
        gonly = _with_host_plan(host, fit)
        assert gonly.fit_stats()["root_gradient_only_builds"] == 5
        assert gonly.serialize() == packed.serialize()
        for k in ("histogram_rows", "histogram_bytes", "histogram_launches", "trees", "resolved_nodes"):
            assert gonly.fit_stats()[k] == packed.fit_stats()[k], k


@pytest.mark.parametrize("case", ["weighted_int16", "fp64_fixed", "logistic_int8", "softmax_int16",
                                  "softmax_batched", "fp32"])
def test_count_free_roots_match_packed_roots(case):
    # Every integer form builds its roots after a fit's first without their count plane ([G|H]
    # words, or fixed-point G): model bytes equal the packed roots' (MPSB_ROOT_GONLY=0) under
    # both planners. The FP32 form keeps packed roots.
    X, y = _data(rows=24_000, feats=20)
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, 64)
    w, classes, kw = None, 0, {}
    if case == "weighted_int16":
        prec, w = "int16", np.linspace(0.5, 2.0, X.shape[0])
    elif case == "fp64_fixed":
        prec = "fp64_fixed"
    elif case == "logistic_int8":
        prec, y, kw = "int8", (y > np.median(y)).astype(np.float64), {"objective": "binary_logistic"}
    elif case.startswith("softmax"):
        prec, classes = "int16", 4
        y = np.digitize(y, np.quantile(y, [0.25, 0.5, 0.75])).astype(np.float64)
        if case == "softmax_batched":
            kw = {"class_rounds": "batched", "near_ties": "keep"}
    else:
        prec = "fp32"
    p = N._TrainingParameters(4, 0.3, 64, 5, 5, 0.0, 1.0, precision=prec, rounding_seed=3, **kw)
    fit = lambda: N._train_binned(s, ds, y, w, p, classes)  # noqa: E731
    packed = _with_env("MPSB_ROOT_GONLY", "0", lambda: _with_host_plan(False, fit))
    assert packed.fit_stats()["root_gradient_only_builds"] == 0
    roots = 4 * max(classes, 1) if case != "softmax_batched" else 4
    planners = (False,) if case in ("softmax_batched", "fp32") else (False, True)
    for host in planners:
        m = _with_host_plan(host, fit)
        free = m.fit_stats()["root_gradient_only_builds"]
        if case == "fp32":
            assert free == 0
            continue
        assert free == roots - 1, (case, host, free)
        if prec == "fp32":
            continue
        assert m.serialize() == packed.serialize(), (case, host)
