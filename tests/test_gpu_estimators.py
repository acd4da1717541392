"""Estimators with device="hip" against the oracle trained with the same parameters
(src/mpsboost/estimators/: base.py, classification.py, multiclass.py, single_tree.py, forest*.py).

Bars: precision="fp64" reproduces the reference's CPU trainer, so squared-error and quantile
predictions equal the oracle's exactly (as float32); objectives that call exp match to rel 1e-6
in probability / prediction space (the device's exp and the host libm's may differ in the last
bit); forest trees equal the oracle's trees fitted on the same bootstrap rows and feature
subsets. Save / load round trips reproduce predictions exactly.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")
import mpsboost_amd as M  # noqa: E402
from mpsboost_amd import estimators as E  # noqa: E402

KW = dict(n_estimators=5, learning_rate=0.3, max_depth=4, min_samples_leaf=20, min_child_weight=1.0)


def _data(seed, rows=4000, feats=6):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    X[rng.random(X.shape) < 0.03] = np.nan
    z = np.sin(np.nan_to_num(X[:, 0])) + 0.5 * np.nan_to_num(X[:, 1]) * np.nan_to_num(X[:, 2])
    return X, z, rng


def _oracle_predict(X, y, w=None, **kw):
    o = O.oracle()
    ds = o.quantize(X, 256)
    return o.train_regressor(ds, y, w, **kw).predict(ds)


@pytest.mark.parametrize("loss,extra", [("squared_error", {}), ("quantile", dict(quantile_alpha=0.7)),
                                        ("poisson", {}), ("tweedie", dict(tweedie_variance_power=1.4))])
def test_regressor_matches_oracle(loss, extra):
    X, z, rng = _data(1)
    y = z + 0.2 * rng.standard_normal(z.size) if loss in ("squared_error", "quantile") else rng.poisson(np.exp(z))
    y = np.asarray(y, dtype=np.float64)
    est = M.GradientBoostingRegressor(loss=loss, precision="fp64", device="hip", **KW, **extra).fit(X, y)
    oracle_kw = dict(objective=loss, **KW)
    if loss == "quantile":
        oracle_kw["objective_alpha"] = extra["quantile_alpha"]
    if loss == "tweedie":
        oracle_kw["tweedie_variance_power"] = extra["tweedie_variance_power"]
    raw = _oracle_predict(X, y, **oracle_kw)
    pred = est.predict(X)
    if loss in ("squared_error", "quantile"):
        assert np.array_equal(pred, raw.astype(np.float32))
    else:
        np.testing.assert_allclose(pred, np.exp(raw).astype(np.float32), rtol=1e-6)
    assert est.device_ == "hip" and est.n_estimators_ == 5
    assert est.score(X, y) <= 1.0
    imp = est.feature_importance()
    assert imp.shape == (6,) and imp.sum() == pytest.approx(1.0, rel=1e-6)


def test_binary_classifier_matches_oracle():
    X, z, rng = _data(2)
    y = (rng.random(z.size) < 1 / (1 + np.exp(-2 * z))).astype(np.int64) * 3 + 4  # labels {4, 7}
    clf = M.GradientBoostingClassifier(precision="fp64", **KW).fit(X, y)
    assert list(clf.classes_) == [4, 7]
    margins = _oracle_predict(X, (y == 7).astype(np.float64), objective="binary_logistic", **KW)
    p = 1 / (1 + np.exp(-margins))
    np.testing.assert_allclose(clf.predict_proba(X)[:, 1], p.astype(np.float32), rtol=1e-6, atol=1e-7)
    assert set(np.unique(clf.predict(X))) <= {4, 7}
    assert 0.5 < clf.score(X, y) <= 1.0


@pytest.mark.parametrize("strategy", ["softmax", "ovr"])
def test_multiclass_classifier(strategy):
    X, z, rng = _data(3)
    y = np.digitize(z + 0.3 * rng.standard_normal(z.size), [-0.5, 0.3, 1.0]) * 10.0  # classes 0, 10, 20, 30
    clf = M.GradientBoostingClassifier(precision="fp64", multi_strategy=strategy, **KW).fit(X, y)
    assert list(clf.classes_) == [0.0, 10.0, 20.0, 30.0]
    proba = clf.predict_proba(X)
    assert proba.shape == (X.shape[0], 4)
    np.testing.assert_allclose(proba.sum(axis=1), 1.0, rtol=1e-5)
    if strategy == "softmax":
        o = O.oracle()
        ds = o.quantize(X, 256)
        ref = o.train_multiclass(ds, np.searchsorted(clf.classes_, y).astype(np.float64), 4, **KW)
        margins = ref.predict(ds).reshape(-1, 4)
        e = np.exp(margins - margins.max(axis=1, keepdims=True))
        np.testing.assert_allclose(proba, (e / e.sum(axis=1, keepdims=True)).astype(np.float32), rtol=1e-5,
                                   atol=1e-6)
    assert clf.score(X, y) > 0.4


def test_decision_and_extra_trees():
    X, z, rng = _data(4)
    y = z + 0.1 * rng.standard_normal(z.size)
    t = M.DecisionTreeRegressor(max_depth=5, precision="fp64").fit(X, y)
    raw = _oracle_predict(X, y, n_estimators=1, learning_rate=1.0, max_depth=5, min_samples_leaf=20,
                          min_child_weight=1.0)
    assert np.array_equal(t.predict(X), raw.astype(np.float32))
    x = M.ExtraTreeRegressor(max_depth=5, random_state=11, precision="fp64").fit(X, y)
    raw = _oracle_predict(X, y, n_estimators=1, learning_rate=1.0, max_depth=5, min_samples_leaf=20,
                          min_child_weight=1.0, split_strategy="random_threshold", random_seed=11)
    assert np.array_equal(x.predict(X), raw.astype(np.float32))


def test_random_forest_classifier_trees_match_oracle():
    """Per-tree seeds, bootstrap rows and feature subsets as forest.py:120-129; each tree equals
    the oracle's one-tree binary-logistic fit on the same sample (lr 1, fp64 form)."""

    X, z, rng = _data(5, rows=3000, feats=8)
    y = (z > 0.2).astype(np.float64)
    forest = M.RandomForestClassifier(n_estimators=4, max_depth=5, max_features=0.5, random_state=42,
                                      precision="fp64").fit(X, y)
    gen = np.random.default_rng(42)
    o = O.oracle()
    for tree, feats in zip(forest.estimators_, forest.feature_subsets_):
        row_seed, feat_seed = (int(gen.integers(0, np.iinfo(np.int32).max)) for _ in range(2))
        rows = None
        for attempt in range(16):
            cand = E.bootstrap_sample_indices(3000, random_state=row_seed + attempt)
            if np.unique(y[cand]).size == 2:
                rows = cand
                break
        assert np.array_equal(feats, E.subsample_feature_indices(8, feature_fraction=0.5, random_state=feat_seed))
        Xs = X[rows][:, feats]
        ds = o.quantize(Xs, 256)
        ref = o.train_regressor(ds, y[rows], np.ones(rows.size), n_estimators=1, learning_rate=1.0,
                                objective="binary_logistic", max_depth=5, min_samples_leaf=20, min_child_weight=1.0,
                                random_seed=feat_seed)
        a, b = tree.model_.trees[0], ref.trees[0]
        for k in ("feature_index", "threshold_bin", "left_child", "right_child", "default_left", "flags"):
            assert np.array_equal(a[k], b[k]), k
        np.testing.assert_allclose(a["leaf_value"], b["leaf_value"], rtol=1e-9, atol=1e-12)
    proba = forest.predict_proba(X)
    assert proba.shape == (3000, 2)
    np.testing.assert_allclose(proba.sum(axis=1), 1.0, rtol=1e-6)
    assert forest.score(X, y) > 0.6
    imp = forest.feature_importance()
    assert imp.shape == (8,) and imp.sum() == pytest.approx(1.0, rel=1e-6)


def test_random_forest_regressor_and_extra_trees():
    X, z, rng = _data(6, rows=2500)
    y = z + 0.1 * rng.standard_normal(z.size)
    for cls in (M.RandomForestRegressor, M.ExtraTreesRegressor):
        f = cls(n_estimators=3, max_depth=4, sample_fraction=0.5, bootstrap=False, random_state=3).fit(X, y)
        assert f.predict(X).shape == (2500,)
        assert f.score(X, y) > 0.3
    c = M.ExtraTreesClassifier(n_estimators=3, max_depth=4, random_state=3).fit(X, (z > 0).astype(np.float64))
    assert c.predict(X).shape == (2500,)


def test_save_load_round_trip(tmp_path):
    X, z, rng = _data(7)
    y = z + 0.1 * rng.standard_normal(z.size)
    est = M.GradientBoostingRegressor(precision="int16", **KW).fit(X, y)
    path = tmp_path / "m.mpsb"
    est.save_model(str(path))
    loaded = M.GradientBoostingRegressor().load_model(str(path))
    assert np.array_equal(loaded.predict(X), est.predict(X))
    yc = np.digitize(z, [-0.5, 0.5]).astype(np.float64)
    clf = M.GradientBoostingClassifier(**KW).fit(X, yc)
    clf.save_model(str(tmp_path / "c.mpsb"))
    back = M.GradientBoostingClassifier().load_model(str(tmp_path / "c.mpsb"))
    assert back._multiclass_strategy_ == "native_softmax"
    assert np.array_equal(back.predict_proba(X), clf.predict_proba(X))
    assert np.array_equal(back.classes_, clf.classes_)


def test_auto_device_and_summary():
    X, z, rng = _data(8, rows=1000)
    est = M.GradientBoostingRegressor(n_estimators=3, device="auto").fit(X, z)
    assert est.device_ == "hip" and est.training_summary_["precision"] == "fp64_fixed"
    assert est.device_decision_["requested"] == "auto"
