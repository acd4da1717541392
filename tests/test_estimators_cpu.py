"""Host-side behaviour of the estimator surface (src/mpsboost/estimators/, device_policy.py):
device choice, parameter validation, not-fitted errors, sampling helpers and model loading.
No device is used: these run where the GPU is absent."""

import numpy as np
import pytest

from oracle import oracle as O

N = pytest.importorskip("mpsboost_amd._native")
import mpsboost_amd as M  # noqa: E402
from mpsboost_amd import estimators as E  # noqa: E402

REF = O.reference()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref/libref.so was not built")


# The reference's choose_device (src/mpsboost/device_policy.py:52-99) restated as a table of
# (requested, work vs threshold, device available) -> selected, with a CPU trainer present.
@pytest.mark.parametrize("requested,rows,available,selected,reason", [
    ("cpu", 10, True, "cpu", "explicit cpu request"),
    ("hip", 10, False, "hip", "explicit hip request"),
    ("auto", 10, False, "cpu", "hip unavailable"),
    ("auto", 10, True, "cpu", "estimated work below hip threshold"),
    ("auto", 1_000_000, True, "hip", "estimated work meets hip threshold"),
])
def test_choose_device_with_cpu_trainer(requested, rows, available, selected, reason):
    d = M.choose_device(requested=requested, n_samples=rows, n_features=8, n_estimators=100, max_bins=256,
                        hip_available=available, cpu_available=True)
    assert (d.selected, d.reason) == (selected, reason)
    assert d.estimated_work == rows * 8 * 100 * 256


def test_choose_device_without_cpu_trainer_and_errors():
    d = M.choose_device(requested="auto", n_samples=10, n_features=8, n_estimators=1, max_bins=2, hip_available=True)
    assert d.selected == "hip"
    with pytest.raises(ValueError, match="requested device"):
        M.choose_device(requested="mps", n_samples=1, n_features=1, n_estimators=1, max_bins=2, hip_available=True)
    with pytest.raises(TypeError, match="n_samples"):
        M.choose_device(requested="auto", n_samples=1.5, n_features=1, n_estimators=1, max_bins=2,
                        hip_available=True)
    with pytest.raises(ValueError, match="positive"):
        M.choose_device(requested="auto", n_samples=0, n_features=1, n_estimators=1, max_bins=2, hip_available=True)
    with pytest.raises(TypeError, match="hip_available"):
        M.choose_device(requested="auto", n_samples=1, n_features=1, n_estimators=1, max_bins=2, hip_available=1)


@pytest.mark.parametrize("params,error,match", [
    (dict(n_estimators=0), ValueError, r"n_estimators must be in \[1"),
    (dict(n_estimators=2.0), TypeError, "n_estimators must be an integer"),
    (dict(max_depth=32), ValueError, "max_depth"),
    (dict(max_bins=1), ValueError, "max_bins"),
    (dict(learning_rate=0.0), ValueError, r"learning_rate must be in \(0.0, 1.0\]"),
    (dict(reg_lambda=-1.0), ValueError, "reg_lambda"),
    (dict(loss="huber"), ValueError, "loss must be"),
    (dict(quantile_alpha=1.0), ValueError, "quantile_alpha"),
    (dict(tweedie_variance_power=2.0), ValueError, "tweedie_variance_power"),
    (dict(device="mps"), ValueError, "device must be 'hip', 'cpu', or 'auto'"),
    (dict(precision="fp16"), ValueError, "precision must be one of"),
    (dict(max_leaves=5, max_active_leaves=7), ValueError, "max_active_leaves must not exceed max_leaves"),
    (dict(n_jobs=0), ValueError, "n_jobs"),
    (dict(random_state=True), TypeError, "random_state"),
    (dict(categorical_features=[0]), ValueError, "categorical_features"),
])
def test_parameter_validation_before_any_device_work(params, error, match):
    X = np.zeros((10, 2), dtype=np.float32)
    y = np.arange(10, dtype=np.float64)
    with pytest.raises(error, match=match):
        M.GradientBoostingRegressor(**params).fit(X, y)


def test_params_protocol_and_not_fitted():
    r = M.GradientBoostingRegressor(n_estimators=7)
    p = r.get_params()
    assert p["n_estimators"] == 7 and p["device"] == "hip" and p["min_samples_leaf"] == 20
    assert M.GradientBoostingRegressor(**p).get_params() == p
    with pytest.raises(ValueError, match="Unknown parameter"):
        r.set_params(nonsense=1)
    with pytest.raises(M.NotFittedError, match="not fitted"):
        r.predict(np.zeros((2, 2)))
    c = M.GradientBoostingClassifier()
    assert c.get_params()["multi_strategy"] == "auto"
    with pytest.raises(ValueError, match="at least two classes"):
        c.fit(np.zeros((4, 2)), np.zeros(4))
    with pytest.raises(ValueError, match="exactly 0 and 1"):
        M.DecisionTreeClassifier().fit(np.zeros((4, 2)), np.array([0.0, 1.0, 2.0, 1.0]))
    f = M.RandomForestClassifier(max_features=1.5)
    with pytest.raises(ValueError, match=r"max_features must be in \(0, 1\]"):
        f.fit(np.zeros((4, 2)), np.array([0.0, 1.0, 0.0, 1.0]))


def test_no_silent_cpu_path():
    X = np.random.default_rng(0).standard_normal((50, 3)).astype(np.float32)
    y = X[:, 0].astype(np.float64)
    with pytest.raises(M.BackendError, match="CPU trainer"):
        M.GradientBoostingRegressor(n_estimators=2, device="cpu").fit(X, y)
    if not M.is_available():
        for device in ("hip", "auto"):
            with pytest.raises(M.BackendError, match="No supported HIP device"):
                M.GradientBoostingRegressor(n_estimators=2, device=device).fit(X, y)


def test_input_adapters():
    with pytest.raises(ValueError, match="two-dimensional"):
        E.as_dense_matrix(np.zeros(3))
    assert E.as_dense_matrix(np.zeros((2, 2), dtype=np.int64)).dtype == np.float32
    assert E.as_dense_matrix(np.zeros((2, 2))[::-1]).strides[0] > 0
    with pytest.raises(ValueError, match="sample counts"):
        E.as_labels(np.zeros(3), 4)
    with pytest.raises(ValueError, match="non-negative"):
        E.as_sample_weight(np.array([1.0, -1.0]), 2)
    assert np.array_equal(E.as_sample_weight(None, 3), np.ones(3))


def test_sampling_helpers_follow_the_numpy_generator():
    """src/mpsboost/randomization.py:43-80: one default_rng per call, ceil-rounded counts."""

    rows = E.bootstrap_sample_indices(1000, sample_fraction=0.333, random_state=5)
    assert rows.size == 333 and rows.dtype == np.int64 and rows.min() >= 0 and rows.max() < 1000
    assert np.array_equal(rows, E.bootstrap_sample_indices(1000, sample_fraction=0.333, random_state=5))
    unique = E.sample_without_replacement_indices(50, sample_fraction=0.5, random_state=3)
    assert unique.size == 25 and np.unique(unique).size == 25
    feats = E.subsample_feature_indices(64, feature_fraction=0.125, random_state=9)
    assert feats.size == 8 and np.unique(feats).size == 8
    assert E.subsample_feature_indices(3, feature_fraction=0.01, random_state=1).size == 1
    with pytest.raises(ValueError, match="finite and positive"):
        E.bootstrap_sample_indices(10, sample_fraction=0.0)


def _reference_file(tmp_path, objective="squared_error", **kw):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((400, 4)).astype(np.float32)
    y = np.abs(X[:, 0]) + 0.5 if objective != "binary_logistic" else (X[:, 0] > 0).astype(np.float64)
    blob = REF.train_regressor(REF.quantize(X, 16), y, n_estimators=3, learning_rate=0.3, objective=objective,
                               max_depth=3, min_samples_leaf=5, **kw).serialize()
    path = tmp_path / f"{objective}.mpsb"
    path.write_bytes(blob)
    return path


@needs_ref
def test_estimators_load_reference_files(tmp_path):
    path = _reference_file(tmp_path)
    r = M.GradientBoostingRegressor().load_model(str(path))
    assert (r.n_features_in_, r.n_estimators_, r.training_summary_["loaded"]) == (4, 3, True)
    assert r.feature_importance("split").sum() == pytest.approx(1.0)
    with pytest.raises(ValueError, match="incompatible with MPSBoostRegressor"):
        M.GradientBoostingRegressor(loss="poisson").load_model(str(path))
    q = _reference_file(tmp_path, "quantile", objective_alpha=0.3)
    with pytest.raises(ValueError, match="quantile alpha is incompatible"):
        M.GradientBoostingRegressor(loss="quantile", quantile_alpha=0.5).load_model(str(q))
    b = _reference_file(tmp_path, "binary_logistic")
    c = M.GradientBoostingClassifier().load_model(str(b))  # not softmax: the binary loader takes it
    assert c._multiclass_strategy_ == "binary_logistic" and list(c.classes_) == [0, 1]
    with pytest.raises(M.TrainingError, match="multiclass loader"):
        N._load_model(str(path), 1)
    with pytest.raises(M.TrainingError, match="regression loader"):
        N._deserialize_model(REF.train_multiclass(REF.quantize(np.eye(6, dtype=np.float32), 4),
                                                  np.array([0.0, 1, 2, 0, 1, 2]), 3, n_estimators=1,
                                                  max_depth=1).serialize(), 0)
