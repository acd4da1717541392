"""Estimators over the HIP device: the reference's scikit-learn-style surface
(src/mpsboost/estimators/) for the callers of the histogram-boosting path.

Same class names, constructor parameters, defaults, validation messages, label encodings,
sampling seeds and prediction aggregation as the reference:
* ``MPSBoostRegressor`` / ``GradientBoostingRegressor`` (estimators/base.py, model_state.py):
  losses squared_error / quantile / poisson / tweedie, exp link for Poisson and Tweedie;
* ``MPSBoostClassifier`` / ``GradientBoostingClassifier`` (classification.py, multiclass.py):
  binary logistic on two classes; softmax (on the device) or one-vs-rest on more;
* ``DecisionTree*`` / ``ExtraTree*`` (single_tree.py): one tree, learning rate 1;
* ``RandomForest*`` / ``ExtraTrees*`` (forest.py, forest_sampling.py, forest_public.py): the
  reference's per-tree numpy seeds, bootstrap / without-replacement rows, feature subsets and
  mean aggregation; each tree is binned and trained on the device.

Differences, each explicit: ``device`` takes "hip", "cpu" or "auto" ("cpu" names the
reference's CPU trainer, which is not part of this package, and raises); ``precision`` selects
the device's histogram arithmetic (default "fp64_fixed": exact integer sums; "fp64" reproduces
the reference's CPU trainer bit for bit; "int16" / "int8" are the fastest); categorical
encoding is not provided. Nothing falls back to a CPU path: without a device, fit raises
``BackendError``.
"""

from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from threading import Lock
from time import perf_counter
from typing import Any

import numpy as np

from . import _native
from .device_policy import choose_device, decision_to_dict, is_available, setup_instructions

PRECISIONS = ("fp64", "fp64_fixed", "fp32", "int16", "int8")


class NotFittedError(RuntimeError):
    """Fitted-only functionality called before a complete model exists (estimators/errors.py)."""


# ----------------------------------------------------------------- input adapters --
# The input contract of the reference's src/mpsboost/matrix.py:17-89 (which inputs are
# accepted, the conversions, and the error types and messages), stated as data below.
_REAL_KINDS = frozenset("iuf")  # signed, unsigned and floating; bool, complex, object rejected
_KEPT_FLOATS = (np.dtype(np.float32), np.dtype(np.float64))


class _VectorContract:
    """A one-dimensional per-row input: its messages, in the order the checks apply."""

    def __init__(self, not_1d: str, wrong_length: str, not_real: str):
        self.messages = (not_1d, wrong_length, not_real)

    def adapt(self, value: Any, rows: int) -> np.ndarray:
        array = np.asarray(value)
        failures = ((array.ndim != 1, ValueError), (array.ndim == 1 and array.shape[0] != rows, ValueError),
                    (array.dtype.kind not in _REAL_KINDS, TypeError))
        for (failed, kind), message in zip(failures, self.messages):
            if failed:
                raise kind(message)
        return np.ascontiguousarray(array, dtype=np.float64)


_LABELS = _VectorContract("y must be a one-dimensional numeric array", "X and y sample counts do not match",
                          "y dtype must be a real numeric type")
_WEIGHTS = _VectorContract("sample_weight must be a one-dimensional numeric array",
                           "sample_weight length must match X rows", "sample_weight dtype must be numeric")


def as_dense_matrix(value: Any) -> np.ndarray:
    """2-D real input for the device binning: float32 / float64 kept as they are (views
    included), other real types converted to float32, non-positive strides copied to a
    C-contiguous array (the native layer reads positive strides only)."""

    if not isinstance(value, np.ndarray) and hasattr(value, "toarray"):
        raise TypeError("Sparse matrices are not supported")
    array = np.asarray(value)
    if array.ndim != 2:
        raise ValueError("X must be a two-dimensional dense numeric array")
    if 0 in array.shape:
        raise ValueError("X must contain at least one row and one feature")
    if array.dtype.kind not in _REAL_KINDS:
        raise TypeError("X dtype must be a real numeric type")
    if array.dtype not in _KEPT_FLOATS:
        array = array.astype(np.float32)
    return array if min(array.strides) > 0 else np.ascontiguousarray(array)


def as_labels(value: Any, expected_rows: int) -> np.ndarray:
    """Labels as contiguous float64, one per row."""

    return _LABELS.adapt(value, expected_rows)


def as_sample_weight(value: Any, expected_rows: int) -> np.ndarray:
    """Sample weights as contiguous float64: unit weights when None; finite, non-negative,
    with a positive total."""

    if value is None:
        return np.ones(expected_rows, dtype=np.float64)
    weights = _WEIGHTS.adapt(value, expected_rows)
    problems = (("sample_weight values must be finite", lambda w: not np.isfinite(w).all()),
                ("sample_weight values must be non-negative", lambda w: bool((w < 0.0).any())),
                ("sample_weight total must be positive", lambda w: float(w.sum()) <= 0.0))
    for message, found in problems:
        if found(weights):
            raise ValueError(message)
    return weights


# ------------------------------------------------------------------ randomisation --
def _rng(random_state: int | None) -> np.random.Generator:
    if random_state is not None and (isinstance(random_state, bool) or not isinstance(random_state, int)):
        raise TypeError("random_state must be an integer or None")
    return np.random.default_rng(random_state)


def _positive_count(total: int, fraction: float, *, name: str) -> int:
    if isinstance(total, bool) or not isinstance(total, int):
        raise TypeError(f"{name} total must be an integer")
    if total <= 0:
        raise ValueError(f"{name} total must be positive")
    if isinstance(fraction, bool) or not isinstance(fraction, (int, float)):
        raise TypeError(f"{name} fraction must be numeric")
    numeric = float(fraction)
    if not np.isfinite(numeric) or numeric <= 0.0:
        raise ValueError(f"{name} fraction must be finite and positive")
    return max(1, int(math.ceil(total * numeric)))


def bootstrap_sample_indices(n_samples: int, *, sample_fraction: float = 1.0,
                             random_state: int | None = None) -> np.ndarray:
    """src/mpsboost/randomization.py:43-52: rows with replacement."""

    count = _positive_count(n_samples, sample_fraction, name="sample")
    return _rng(random_state).integers(0, n_samples, size=count, dtype=np.int64)


def sample_without_replacement_indices(n_samples: int, *, sample_fraction: float = 1.0,
                                       random_state: int | None = None) -> np.ndarray:
    """src/mpsboost/randomization.py:55-66."""

    count = _positive_count(n_samples, sample_fraction, name="sample")
    if count > n_samples:
        raise ValueError("sample_fraction selects more rows than available without replacement")
    return _rng(random_state).choice(n_samples, size=count, replace=False).astype(np.int64)


def subsample_feature_indices(n_features: int, *, feature_fraction: float = 1.0,
                              random_state: int | None = None) -> np.ndarray:
    """src/mpsboost/randomization.py:69-80."""

    count = _positive_count(n_features, feature_fraction, name="feature")
    if count > n_features:
        raise ValueError("feature_fraction selects more features than available")
    return _rng(random_state).choice(n_features, size=count, replace=False).astype(np.int64)


# --------------------------------------------------------------- parameter checks --
# The reference's accepted ranges, error types and messages for estimator parameters
# (src/mpsboost/estimators/model_state.py:148-274, forest_sampling.py:118-150), as a table
# of checkers; the estimators walk the table in order, so the first bad parameter raises.
class _IntIn:
    """An integer (bool excluded) in [lo, hi]; None passes when optional."""

    def __init__(self, lo: int, hi: int, optional: bool = False, range_message: str | None = None):
        self.lo, self.hi, self.optional, self.range_message = lo, hi, optional, range_message

    def __call__(self, name: str, value: Any) -> None:
        if value is None and self.optional:
            return
        if type(value) is bool or not isinstance(value, int):
            raise TypeError(f"{name} must be an integer")
        if value < self.lo or value > self.hi:
            raise ValueError(self.range_message or f"{name} must be in [{self.lo}, {self.hi}]")


class _RealIn:
    """A finite int or float (bool excluded) in an interval whose ends are open or closed."""

    def __init__(self, lo: float, hi: float, closed: str, kind: str = "a real number", shown: str | None = None):
        self.lo, self.hi, self.kind = lo, hi, kind
        self.lo_closed, self.hi_closed = closed[0] == "[", closed[1] == "]"
        self.shown = shown or f"{closed[0]}{lo}, {hi}{closed[1]}"

    def __call__(self, name: str, value: Any) -> None:
        if type(value) is bool or not isinstance(value, (int, float)):
            raise TypeError(f"{name} must be {self.kind}")
        x = float(value)
        above = x >= self.lo if self.lo_closed else x > self.lo
        below = x <= self.hi if self.hi_closed else x < self.hi
        if not (math.isfinite(x) and above and below):
            raise ValueError(f"{name} must be in {self.shown}")


_UNBOUNDED = float("inf")
_BOOSTING_CHECKS = (
    ("n_estimators", _IntIn(1, 2**32 - 1)),
    ("max_depth", _IntIn(0, 31)),
    ("max_bins", _IntIn(2, 65536)),
    ("min_samples_leaf", _IntIn(1, 2**63 - 1)),
    ("verbosity", _IntIn(0, 3)),
    ("max_leaves", _IntIn(2, 2**32 - 1, optional=True)),
    ("max_active_leaves", _IntIn(2, 2**32 - 1, optional=True)),
    ("n_jobs", _IntIn(-1, 2**31 - 1, optional=True)),
    ("n_jobs", lambda name, v: _raise(ValueError, "n_jobs must be None, -1, or a non-zero integer") if v == 0 else None),
    ("learning_rate", _RealIn(0.0, 1.0, "(]")),
    ("min_child_weight", _RealIn(0.0, _UNBOUNDED, "[]")),
    ("min_gain_to_split", _RealIn(0.0, _UNBOUNDED, "[]")),
    ("reg_lambda", _RealIn(0.0, _UNBOUNDED, "[]")),
    ("reg_alpha", _RealIn(0.0, _UNBOUNDED, "[]")),
    ("max_delta_step", _RealIn(0.0, _UNBOUNDED, "[]")),
)
_OBJECTIVE_CHECKS = (
    ("quantile_alpha", _RealIn(0.0, 1.0, "()", kind="numeric")),
    ("tweedie_variance_power", _RealIn(1.0, 2.0, "()", kind="numeric")),
)
_FOREST_CHECKS = (
    ("n_estimators", _IntIn(1, 2**63 - 1, range_message="n_estimators must be positive")),
    ("n_jobs", _IntIn(1, 2**63 - 1, range_message="n_jobs must be positive")),
    ("bootstrap", lambda name, v: None if type(v) is bool else _raise(TypeError, "bootstrap must be a boolean")),
    ("max_features", _RealIn(0.0, 1.0, "(]", kind="numeric", shown="(0, 1]")),
    ("sample_fraction", _RealIn(0.0, 1.0, "(]", kind="numeric", shown="(0, 1]")),
)
_REGRESSION_LOSSES = ("squared_error", "quantile", "poisson", "tweedie")


def _raise(kind: type, message: str) -> None:
    raise kind(message)


def _run_checks(owner: Any, checks) -> None:
    for name, check in checks:
        check(name, getattr(owner, name))


def _feature_integer(value: Any, what: str) -> int:
    if type(value) is bool or not isinstance(value, (int, np.integer)):
        raise TypeError(f"{what} values must be integers")
    return int(value)


# ------------------------------------------------------------------- regressor -----
class MPSBoostRegressor:
    """Histogram GBDT regressor trained on the HIP device (estimators/base.py:25-318)."""

    _estimator_type = "regressor"
    _native_objective = "squared_error"
    _split_strategy = "best_gain"
    _fitted_error_message = "MPSBoostRegressor is not fitted or loaded"
    _PARAMETER_NAMES = (
        "n_estimators", "learning_rate", "max_depth", "max_bins", "growth_strategy", "max_leaves",
        "max_active_leaves", "min_gain_to_split", "loss", "quantile_alpha", "tweedie_variance_power",
        "min_child_weight", "min_samples_leaf", "reg_lambda", "reg_alpha", "max_delta_step",
        "monotonic_constraints", "interaction_constraints", "categorical_features", "random_state", "n_jobs",
        "device", "verbosity", "precision", "near_ties",
    )

    def __init__(self, n_estimators: int = 100, learning_rate: float = 0.1, max_depth: int = 6, max_bins: int = 256,
                 growth_strategy: str = "level_wise", max_leaves: int | None = None,
                 max_active_leaves: int | None = None, min_gain_to_split: float = 0.0, loss: str = "squared_error",
                 quantile_alpha: float = 0.5, tweedie_variance_power: float = 1.5, min_child_weight: float = 1.0,
                 min_samples_leaf: int = 20, reg_lambda: float = 1.0, reg_alpha: float = 0.0,
                 max_delta_step: float = 0.0, monotonic_constraints: Any = None, interaction_constraints: Any = None,
                 categorical_features: Any = None, random_state: int | None = None, n_jobs: int | None = None,
                 device: str = "hip", verbosity: int = 1, precision: str = "fp64_fixed",
                 near_ties: str = "auto") -> None:
        self.n_estimators = n_estimators
        self.learning_rate = learning_rate
        self.max_depth = max_depth
        self.max_bins = max_bins
        self.growth_strategy = growth_strategy
        self.max_leaves = max_leaves
        self.max_active_leaves = max_active_leaves
        self.min_gain_to_split = min_gain_to_split
        self.loss = loss
        self.quantile_alpha = quantile_alpha
        self.tweedie_variance_power = tweedie_variance_power
        self.min_child_weight = min_child_weight
        self.min_samples_leaf = min_samples_leaf
        self.reg_lambda = reg_lambda
        self.reg_alpha = reg_alpha
        self.max_delta_step = max_delta_step
        self.monotonic_constraints = monotonic_constraints
        self.interaction_constraints = interaction_constraints
        self.categorical_features = categorical_features
        self.random_state = random_state
        self.n_jobs = n_jobs
        self.device = device
        self.verbosity = verbosity
        self.precision = precision
        self.near_ties = near_ties
        self._fit_lock = Lock()

    # -- parameters
    def get_params(self, deep: bool = True) -> dict[str, Any]:
        del deep
        return {name: getattr(self, name) for name in self._PARAMETER_NAMES}

    def set_params(self, **parameters: Any):
        unknown = sorted(set(parameters) - set(self._PARAMETER_NAMES))
        if unknown:
            raise ValueError(f"Unknown parameter(s): {', '.join(unknown)}")
        for name, value in parameters.items():
            setattr(self, name, value)
        if parameters:
            self._clear_fitted_state()
        return self

    def _validate_parameters(self) -> None:
        """The reference's parameter contract (estimators/model_state.py:148-203), with the
        device values of this package: the checks table first, then the option choices."""

        _run_checks(self, _BOOSTING_CHECKS)
        choices = (
            (self.growth_strategy in ("level_wise", "leaf_wise"), "growth_strategy must be 'level_wise' or 'leaf_wise'"),
            (self.growth_strategy != "leaf_wise",
             "growth_strategy 'leaf_wise' is not available on the hip device (level_wise only)"),
            (None in (self.max_leaves, self.max_active_leaves) or self.max_active_leaves <= self.max_leaves,
             "max_active_leaves must not exceed max_leaves"),
            (self.device in ("hip", "cpu", "auto"), "device must be 'hip', 'cpu', or 'auto'"),
            (self.precision in PRECISIONS, f"precision must be one of {', '.join(PRECISIONS)}"),
            (getattr(self, "near_ties", "auto") in _native.NEAR_TIES,
             f"near_ties must be one of {', '.join(_native.NEAR_TIES)}"),
        )
        for ok, message in choices:
            if not ok:
                raise ValueError(message)
        if self.random_state is not None and (type(self.random_state) is bool or not isinstance(self.random_state, int)):
            raise TypeError("random_state must be an integer or None")
        if self.categorical_features is not None:
            raise ValueError("categorical_features are not supported by the hip estimators")
        self._validate_objective_parameters()

    def _resolved_native_objective(self) -> str:
        if self._native_objective != "squared_error":
            return self._native_objective
        return self.loss

    def _validate_objective_parameters(self) -> None:
        """Loss choice and objective parameters (estimators/model_state.py:212-231): regression
        estimators choose a loss, the classifiers keep squared_error in that slot."""

        loss = getattr(self, "loss", "squared_error")
        if self._native_objective != "squared_error" and loss != "squared_error":
            raise ValueError("loss is only supported by regression estimators")
        if loss not in _REGRESSION_LOSSES:
            raise ValueError("loss must be 'squared_error', 'quantile', 'poisson', or 'tweedie'")
        _run_checks(self, _OBJECTIVE_CHECKS)

    def _normalized_monotonic_constraints(self, n_features: int) -> list[int]:
        """One direction in {-1, 0, 1} per feature, or [] (estimators/model_state.py:233-250)."""

        raw = getattr(self, "monotonic_constraints", None)
        if raw is None:
            return []
        values = list(raw)
        if len(values) != n_features:
            raise ValueError("monotonic_constraints length must match feature count")
        def direction(value: Any) -> int:
            d = _feature_integer(value, "monotonic_constraints")
            if d not in (-1, 0, 1):
                raise ValueError("monotonic_constraints values must be -1, 0, or 1")
            return d

        return [direction(v) for v in values]

    def _normalized_interaction_constraints(self, n_features: int) -> list[list[int]]:
        """Groups of feature indices, negatives counted from the end, each sorted
        (estimators/model_state.py:252-274)."""

        raw = getattr(self, "interaction_constraints", None)
        if raw is None:
            return []

        def index(value: Any) -> int:
            i = _feature_integer(value, "interaction_constraints")
            i = i + n_features if i < 0 else i
            if i < 0 or i >= n_features:
                raise ValueError("interaction constraint feature index is out of range")
            return i

        groups = []
        for group in raw:
            members = [index(v) for v in group]
            if not members:
                raise ValueError("interaction constraint groups must be non-empty")
            if len(members) != len(set(members)):
                raise ValueError("interaction constraint groups must not contain duplicates")
            groups.append(sorted(members))
        return groups

    def _training_parameters(self, n_features: int, objective: str) -> _native._TrainingParameters:
        return _native._TrainingParameters(
            self.n_estimators, self.learning_rate, self.max_bins, self.max_depth, self.min_samples_leaf,
            self.min_child_weight, self.reg_lambda, reg_alpha=self.reg_alpha, max_delta_step=self.max_delta_step,
            min_gain_to_split=self.min_gain_to_split, objective=objective, split_strategy=self._split_strategy,
            growth_strategy=self.growth_strategy, max_leaves=0 if self.max_leaves is None else self.max_leaves,
            max_active_leaves=0 if self.max_active_leaves is None else self.max_active_leaves,
            random_seed=0 if self.random_state is None else self.random_state,
            monotonic_constraints=self._normalized_monotonic_constraints(n_features),
            interaction_constraints=self._normalized_interaction_constraints(n_features),
            objective_alpha=self.quantile_alpha, tweedie_variance_power=self.tweedie_variance_power,
            precision=self.precision, near_ties=getattr(self, "near_ties", "auto"),
        )

    def _decide_device(self, n_samples: int, n_features: int):
        available = is_available()
        decision = choose_device(requested=self.device, n_samples=int(n_samples), n_features=int(n_features),
                                 n_estimators=int(self.n_estimators), max_bins=int(self.max_bins),
                                 hip_available=available)
        if decision.selected == "cpu":
            if self.device == "cpu":
                raise _native.BackendError("device='cpu' selects the reference's CPU trainer, which this package "
                                           "does not include; use device='hip'")
            raise _native.BackendError(setup_instructions())
        if not available:
            raise _native.BackendError(setup_instructions())
        return decision

    # -- training
    def _training_labels(self, y: Any, n_samples: int) -> np.ndarray:
        return as_labels(y, n_samples)

    def _finalize_fitted_metadata(self) -> None:
        pass

    def _validate_loaded_model(self, model: Any) -> None:
        pass

    def _clear_fitted_state(self) -> None:
        for name in ("model_", "n_features_in_", "device_", "device_decision_", "n_estimators_", "training_summary_",
                     "classes_", "estimators_", "feature_subsets_", "_multiclass_strategy_", "_resolved_objective_"):
            self.__dict__.pop(name, None)

    def fit(self, X: Any, y: Any, sample_weight: Any = None):
        """Train a complete model; fitted state is replaced only on success (base.py:142-261)."""

        if not self._fit_lock.acquire(blocking=False):
            raise RuntimeError("Concurrent fit calls on the same estimator are not supported")
        try:
            self._validate_parameters()
            matrix = as_dense_matrix(X)
            labels = self._training_labels(y, matrix.shape[0])
            weights = as_sample_weight(sample_weight, matrix.shape[0])
            objective = self._resolved_native_objective()
            parameters = self._training_parameters(matrix.shape[1], objective)
            decision = self._decide_device(*matrix.shape)
            started = perf_counter()
            candidate = _native._train_regressor_hip(matrix, labels, weights, parameters)
            elapsed = perf_counter() - started
            self._commit(candidate, matrix.shape[1], decision, objective, elapsed, sample_weight is not None,
                         bool(matrix.flags.c_contiguous))
            return self
        except Exception:
            self._clear_fitted_state()
            raise
        finally:
            self._fit_lock.release()

    def _commit(self, candidate, n_features: int, decision, objective: str, elapsed: float, weighted: bool,
                contiguous: bool) -> None:
        """Replace the fitted state with a complete model (base.py:214-258)."""

        self.model_ = candidate
        self.n_features_in_ = n_features
        self.device_ = decision.selected
        self.device_decision_ = decision_to_dict(decision)
        self.n_estimators_ = candidate.tree_count
        self._resolved_objective_ = objective
        self._finalize_fitted_metadata()
        self.training_summary_ = {
            "fit_seconds": elapsed,
            "input_contiguous": contiguous,
            "device": decision.selected,
            "device_decision": self.device_decision_,
            "n_estimators": candidate.tree_count,
            "native_objective": objective,
            "precision": self.precision,
            "weighted": weighted,
            "monotonic_constraints": self._normalized_monotonic_constraints(n_features),
            "interaction_constraints": self._normalized_interaction_constraints(n_features),
            "growth_strategy": self.growth_strategy,
            "max_leaves": self.max_leaves,
            "max_active_leaves": self.max_active_leaves,
            "min_gain_to_split": float(self.min_gain_to_split),
            "loss": self.loss,
            "quantile_alpha": float(self.quantile_alpha),
            "tweedie_variance_power": float(self.tweedie_variance_power),
            "reg_alpha": float(self.reg_alpha),
            "max_delta_step": float(self.max_delta_step),
        }

    def _fit_sampled(self, matrix: "_native._DeviceMatrix", rows: np.ndarray, features: np.ndarray, labels,
                     sample_weight, labels_checked: bool = False):
        """Fit on matrix[rows][:, features] of a device-resident matrix: the same model as
        ``fit(X[rows][:, features], labels, sample_weight)``, without materialising the
        sub-matrix on the host (forest trees). ``labels_checked``: the labels are a sample of
        labels the caller validated for this estimator type (a forest's), so they are used as
        they are; no sample weights then means unit weights."""

        if not self._fit_lock.acquire(blocking=False):
            raise RuntimeError("Concurrent fit calls on the same estimator are not supported")
        try:
            self._validate_parameters()
            unit = labels_checked and sample_weight is None
            y = labels if labels_checked else self._training_labels(labels, int(rows.size))
            w = None if unit else as_sample_weight(sample_weight, int(rows.size))
            objective = self._resolved_native_objective()
            parameters = self._training_parameters(int(features.size), objective)
            decision = self._decide_device(int(rows.size), int(features.size))
            started = perf_counter()
            session = _native._session()
            ds = _native._quantize_matrix_sampled(session, matrix, rows, features, self.max_bins)
            _native._apply_tuning(session, ds, parameters)
            weights = None if unit or np.all(w == 1.0) else w
            # The model keeps its own frozen schema, not the sample's binned codes (which then
            # never leave the device).
            candidate = _native._RegressionModel(_native._train_binned(session, ds, y, weights, parameters), None,
                                                 session)
            self._commit(candidate, int(features.size), decision, objective, perf_counter() - started,
                         sample_weight is not None, True)
            return self
        except Exception:
            self._clear_fitted_state()
            raise
        finally:
            self._fit_lock.release()

    # -- prediction
    def _require_model(self) -> Any:
        if not hasattr(self, "model_"):
            raise NotFittedError(self._fitted_error_message)
        return self.model_

    def _predict_raw(self, X: Any) -> np.ndarray:
        """Raw model output as float32 (estimators/prediction.py:18-33)."""

        model = self._require_model()
        matrix = as_dense_matrix(X)
        if matrix.shape[1] != self.n_features_in_:
            raise ValueError("prediction feature count does not match training data")
        return np.asarray(model.predict(matrix), dtype=np.float32)

    def predict(self, X: Any) -> np.ndarray:
        raw = self._predict_raw(X)
        if getattr(self, "_resolved_objective_", self._native_objective) in {"poisson", "tweedie"}:
            return np.exp(raw.astype(np.float64, copy=False)).astype(np.float32)
        return raw

    def score(self, X: Any, y: Any, sample_weight: Any = None) -> float:
        """Weighted R^2 (base.py:297-313)."""

        predictions = self.predict(X).astype(np.float64, copy=False)
        labels = as_labels(y, predictions.shape[0])
        weights = as_sample_weight(sample_weight, predictions.shape[0])
        residual = float(np.sum(weights * (labels - predictions) ** 2))
        centered = labels - float(np.average(labels, weights=weights))
        total = float(np.sum(weights * centered**2))
        if total == 0.0:
            return 1.0 if residual == 0.0 else 0.0
        value = 1.0 - residual / total
        if not np.isfinite(value):
            raise ValueError("R2 score is not finite")
        return float(value)

    # -- explanation
    @property
    def feature_importances_(self) -> np.ndarray:
        return self.feature_importance(kind="gain")

    def feature_importance(self, kind: str = "gain") -> np.ndarray:
        """Normalised gain sums or split counts of the trained trees (importance.py:17-44)."""

        model = self._require_model()
        if kind not in {"gain", "split"}:
            raise ValueError("feature importance kind must be 'gain' or 'split'")
        values = np.zeros(self.n_features_in_, dtype=np.float64)
        for nodes in model.trees:
            branch = (nodes["flags"] & 1) == 0
            feats = nodes["feature_index"][branch].astype(np.int64)
            if kind == "gain":
                np.add.at(values, feats, np.maximum(nodes["gain"][branch], 0.0))
            else:
                np.add.at(values, feats, 1.0)
        total = float(values.sum())
        if total > 0.0:
            values /= total
        return values.astype(np.float32, copy=False)

    # -- persistence
    def save_model(self, path: str) -> None:
        """MPSBMOD 1.4 file, readable by the reference (model_state.py:55-62)."""

        self._require_model().save(str(path))

    def load_model(self, path: str):
        """model_state.py:64-103: replace fitted state only after complete success."""

        if not self._fit_lock.acquire(blocking=False):
            raise RuntimeError("Model training or loading is already in progress")
        try:
            candidate = _native._load_regression_model(str(path))
            if candidate.objective != self._resolved_native_objective():
                raise ValueError(f"model objective '{candidate.objective}' is incompatible with {type(self).__name__}")
            if candidate.objective == "quantile" and not np.isclose(candidate.objective_alpha,
                                                                     float(self.quantile_alpha)):
                raise ValueError("model quantile alpha is incompatible with estimator")
            if candidate.objective == "tweedie" and not np.isclose(candidate.tweedie_variance_power,
                                                                    float(self.tweedie_variance_power)):
                raise ValueError("model tweedie variance power is incompatible with estimator")
            self._validate_loaded_model(candidate)
            self.model_ = candidate
            self.n_features_in_ = candidate.feature_count
            self.device_ = "hip"
            self.n_estimators_ = candidate.tree_count
            self._resolved_objective_ = candidate.objective
            self._finalize_fitted_metadata()
            self.training_summary_ = {"loaded": True, "device": self.device_, "native_objective": candidate.objective,
                                      "loss": getattr(self, "loss", "squared_error"),
                                      "quantile_alpha": float(getattr(self, "quantile_alpha", 0.5)),
                                      "tweedie_variance_power": float(getattr(self, "tweedie_variance_power", 1.5))}
            return self
        finally:
            self._fit_lock.release()

    def _more_tags(self) -> dict[str, Any]:
        return {"allow_nan": True, "requires_fit": True}


GradientBoostingRegressor = MPSBoostRegressor


# ------------------------------------------------------------------ classifiers ----
def _sigmoid(margins: np.ndarray) -> np.ndarray:
    """Stable logistic (classification.py:45-55)."""

    p = np.empty_like(margins, dtype=np.float64)
    pos = margins >= 0.0
    p[pos] = 1.0 / (1.0 + np.exp(-margins[pos]))
    e = np.exp(margins[~pos])
    p[~pos] = e / (1.0 + e)
    return p


class _BinaryClassifier(MPSBoostRegressor):
    """Binary-logistic classifier (estimators/classification.py)."""

    _estimator_type = "classifier"
    _native_objective = "binary_logistic"
    _fitted_error_message = "MPSBoostClassifier is not fitted or loaded"

    def _training_labels(self, y: Any, n_samples: int) -> np.ndarray:
        labels = as_labels(y, n_samples)
        if not np.all((labels == 0.0) | (labels == 1.0)):
            raise ValueError("binary classification labels must be exactly 0 and 1")
        if np.unique(labels).size != 2:
            raise ValueError("binary classification requires both class 0 and class 1")
        return labels

    def _finalize_fitted_metadata(self) -> None:
        self.classes_ = np.array([0, 1], dtype=np.int64)

    def predict_proba(self, X: Any) -> np.ndarray:
        p = _sigmoid(self._predict_raw(X).astype(np.float64, copy=False))
        return np.column_stack((1.0 - p, p)).astype(np.float32)

    def predict(self, X: Any) -> np.ndarray:
        return self.classes_[(self.predict_proba(X)[:, 1] >= 0.5).astype(np.int64)]

    def score(self, X: Any, y: Any, sample_weight: Any = None) -> float:
        predictions = self.predict(X)
        labels = as_labels(y, predictions.shape[0])
        weights = as_sample_weight(sample_weight, predictions.shape[0])
        return float(np.average((predictions == labels).astype(np.float64), weights=weights))

    def _more_tags(self) -> dict[str, Any]:
        tags = super()._more_tags()
        tags["binary_only"] = True
        return tags


class MPSBoostClassifier(_BinaryClassifier):
    """Binary or multiclass classifier (estimators/multiclass.py): two classes train the binary
    logistic objective; more train the device softmax trainer (multi_strategy "auto" or
    "softmax") or one binary model per class ("ovr")."""

    _PARAMETER_NAMES = MPSBoostRegressor._PARAMETER_NAMES + ("multi_strategy",)

    def __init__(self, n_estimators: int = 100, learning_rate: float = 0.1, max_depth: int = 6, max_bins: int = 256,
                 growth_strategy: str = "level_wise", max_leaves: int | None = None,
                 max_active_leaves: int | None = None, min_gain_to_split: float = 0.0, loss: str = "squared_error",
                 quantile_alpha: float = 0.5, tweedie_variance_power: float = 1.5, min_child_weight: float = 1.0,
                 min_samples_leaf: int = 20, reg_lambda: float = 1.0, reg_alpha: float = 0.0,
                 max_delta_step: float = 0.0, monotonic_constraints: Any = None, interaction_constraints: Any = None,
                 categorical_features: Any = None, random_state: int | None = None, n_jobs: int | None = None,
                 device: str = "hip", verbosity: int = 1, precision: str = "fp64_fixed",
                 multi_strategy: str = "auto", near_ties: str = "auto") -> None:
        super().__init__(n_estimators=n_estimators, learning_rate=learning_rate, max_depth=max_depth,
                         max_bins=max_bins, growth_strategy=growth_strategy, max_leaves=max_leaves,
                         max_active_leaves=max_active_leaves, min_gain_to_split=min_gain_to_split, loss=loss,
                         quantile_alpha=quantile_alpha, tweedie_variance_power=tweedie_variance_power,
                         min_child_weight=min_child_weight, min_samples_leaf=min_samples_leaf,
                         reg_lambda=reg_lambda, reg_alpha=reg_alpha, max_delta_step=max_delta_step,
                         monotonic_constraints=monotonic_constraints,
                         interaction_constraints=interaction_constraints, categorical_features=categorical_features,
                         random_state=random_state, n_jobs=n_jobs, device=device, verbosity=verbosity,
                         precision=precision, near_ties=near_ties)
        self.multi_strategy = multi_strategy

    def fit(self, X: Any, y: Any, sample_weight: Any = None):
        labels = as_labels(y, np.asarray(X).shape[0])
        classes = np.unique(labels)
        if classes.size < 2:
            raise ValueError("classification requires at least two classes")
        # "batched_softmax" is this device's extension (the reference's message lists its own
        # three): softmax rounds whose class trees grow together from the round's margins.
        if self.multi_strategy not in {"auto", "softmax", "ovr", "batched_softmax"}:
            raise ValueError("multi_strategy must be 'auto', 'softmax', or 'ovr'")
        encoded = np.searchsorted(classes, labels)
        if classes.size == 2:
            fitted = super().fit(X, encoded, sample_weight=sample_weight)
            fitted.classes_ = classes
            fitted._multiclass_strategy_ = "binary_logistic"
            return fitted
        if self.multi_strategy == "ovr":
            return self._fit_ovr(X, encoded, classes, sample_weight)
        return self._fit_softmax(X, encoded, classes, sample_weight)

    def _fit_softmax(self, X, encoded, classes, sample_weight):
        """TrainMulticlassSoftmaxModel on the device (multiclass.py:301-370)."""

        if not self._fit_lock.acquire(blocking=False):
            raise RuntimeError("Concurrent fit calls on the same estimator are not supported")
        try:
            self._validate_parameters()
            matrix = as_dense_matrix(X)
            weights = as_sample_weight(sample_weight, matrix.shape[0])
            parameters = self._training_parameters(matrix.shape[1], "squared_error")
            batched = self.multi_strategy == "batched_softmax"
            if batched:
                parameters.class_rounds = "batched"
            strategy = "native_softmax"  # the model is a softmax model either way
            decision = self._decide_device(*matrix.shape)
            started = perf_counter()
            model = _native._train_multiclass_softmax_hip(matrix, encoded.astype(np.float64), weights, parameters,
                                                         int(classes.size))
            model.set_class_labels(classes.astype(np.float64))
            elapsed = perf_counter() - started
            self.model_ = model
            self.classes_ = classes.copy()
            self.n_features_in_ = matrix.shape[1]
            self.device_ = decision.selected
            self.device_decision_ = decision_to_dict(decision)
            self.n_estimators_ = model.tree_count
            self._multiclass_strategy_ = strategy
            self.training_summary_ = {"strategy": strategy, "requested_strategy": self.multi_strategy,
                                      "class_rounds": "batched" if batched else "class_order",
                                      "device": decision.selected, "classes": classes.tolist(),
                                      "n_classes": int(classes.size), "n_estimators": model.tree_count,
                                      "fit_seconds": elapsed, "precision": self.precision,
                                      "weighted": bool(sample_weight is not None)}
            return self
        except Exception:
            self._clear_fitted_state()
            raise
        finally:
            self._fit_lock.release()

    def _fit_ovr(self, X, encoded, classes, sample_weight):
        """One binary model per class (multiclass.py:118-160), trained one after another on the
        device's session."""

        self._validate_parameters()
        weights = as_sample_weight(sample_weight, encoded.shape[0])
        estimators = []
        for index in range(classes.size):
            params = {k: v for k, v in self.get_params().items() if k != "multi_strategy"}
            if self.random_state is not None:
                params["random_state"] = self.random_state + index
            est = _BinaryClassifier(**params)
            est.fit(X, (encoded == index).astype(np.float64), sample_weight=weights)
            estimators.append(est)
        self.classes_ = classes.copy()
        self.estimators_ = tuple(estimators)
        self.n_features_in_ = estimators[0].n_features_in_
        self.n_estimators_ = sum(e.n_estimators_ for e in estimators)
        self.device_ = estimators[0].device_
        self._multiclass_strategy_ = "one_vs_rest"
        self.training_summary_ = {"strategy": "one_vs_rest", "requested_strategy": self.multi_strategy,
                                  "device": self.device_, "classes": classes.tolist(),
                                  "n_classes": int(classes.size), "n_estimators": self.n_estimators_,
                                  "n_jobs": 1, "weighted": bool(sample_weight is not None)}
        return self

    def _softmax_margins(self, X: Any) -> np.ndarray:
        model = self._require_model()
        matrix = as_dense_matrix(X)
        if matrix.shape[1] != self.n_features_in_:
            raise ValueError("prediction feature count does not match training data")
        return np.asarray(model.predict_margins(matrix), dtype=np.float64).reshape(matrix.shape[0], len(self.classes_))

    def _ovr_estimators(self):
        if not hasattr(self, "estimators_"):
            raise NotFittedError(self._fitted_error_message)
        return self.estimators_

    def decision_function(self, X: Any) -> np.ndarray:
        """Binary margins, softmax margins or one-vs-rest margins (multiclass.py:201-212)."""

        strategy = getattr(self, "_multiclass_strategy_", None)
        if strategy == "native_softmax":
            return self._softmax_margins(X).astype(np.float32, copy=False)
        if strategy != "one_vs_rest":
            return self._predict_raw(X).astype(np.float32, copy=False)
        return np.column_stack([e._predict_raw(X) for e in self._ovr_estimators()]).astype(np.float32, copy=False)

    def predict_proba(self, X: Any) -> np.ndarray:
        """multiclass.py:162-181."""

        strategy = getattr(self, "_multiclass_strategy_", None)
        if strategy == "native_softmax":
            margins = self._softmax_margins(X)
            shifted = margins - margins.max(axis=1, keepdims=True)
            p = np.exp(shifted)
            p /= p.sum(axis=1, keepdims=True)
            return p.astype(np.float32, copy=False)
        if strategy != "one_vs_rest":
            return super().predict_proba(X)
        self._ovr_estimators()
        scores = self.decision_function(X)
        scores = 1.0 / (1.0 + np.exp(-np.clip(scores, -709.0, 709.0)))
        sums = scores.sum(axis=1, keepdims=True)
        zero = sums[:, 0] <= 0.0
        if np.any(zero):
            scores[zero, :] = 1.0 / float(len(self.estimators_))
            sums = scores.sum(axis=1, keepdims=True)
        return (scores / sums).astype(np.float32)

    def predict(self, X: Any) -> np.ndarray:
        if getattr(self, "_multiclass_strategy_", None) in {"native_softmax", "one_vs_rest"}:
            return self.classes_[np.argmax(self.predict_proba(X), axis=1)]
        return super().predict(X)

    def score(self, X: Any, y: Any, sample_weight: Any = None) -> float:
        predictions = self.predict(X)
        labels = as_labels(y, predictions.shape[0])
        weights = as_sample_weight(sample_weight, predictions.shape[0])
        return float(np.average((predictions == labels).astype(np.float64), weights=weights))

    def feature_importance(self, kind: str = "gain") -> np.ndarray:
        """multiclass.py:214-236: zeros for softmax models, the OvR mean otherwise."""

        strategy = getattr(self, "_multiclass_strategy_", None)
        if strategy == "native_softmax":
            import warnings

            warnings.warn("Native softmax feature-importance metadata is unavailable in the compact multiclass model "
                          "record; returning zeros. Use permutation_importance(X, y) for model-agnostic multiclass "
                          "explanations.", RuntimeWarning, stacklevel=2)
            return np.zeros(self.n_features_in_, dtype=np.float32)
        if strategy != "one_vs_rest":
            return super().feature_importance(kind=kind)
        values = np.zeros(self.n_features_in_, dtype=np.float64)
        for e in self._ovr_estimators():
            values += e.feature_importance(kind=kind).astype(np.float64, copy=False)
        values /= float(len(self.estimators_))
        return values.astype(np.float32)

    def save_model(self, path: str) -> None:
        if getattr(self, "_multiclass_strategy_", None) == "one_vs_rest":
            raise NotImplementedError("one-vs-rest multiclass model persistence requires native container-format "
                                      "support")
        self._require_model().save(str(path))

    def load_model(self, path: str):
        """Softmax file first, else the binary loader (multiclass.py:255-290)."""

        if not self._fit_lock.acquire(blocking=False):
            raise RuntimeError("Model training or loading is already in progress")
        try:
            try:
                candidate = _native._load_multiclass_model(str(path))
            except ValueError as exc:
                if "multiclass loader" not in str(exc):
                    raise
                candidate = None
        finally:
            self._fit_lock.release()
        if candidate is None:
            self._multiclass_strategy_ = "binary_logistic"
            return super().load_model(path)
        self.model_ = candidate
        self.classes_ = np.asarray(candidate.class_labels, dtype=np.float64)
        self.n_features_in_ = candidate.feature_count
        self.device_ = "hip"
        self.n_estimators_ = candidate.tree_count
        self._multiclass_strategy_ = "native_softmax"
        self.training_summary_ = {"loaded": True, "device": self.device_, "strategy": "native_softmax",
                                  "n_classes": int(candidate.class_count)}
        return self


GradientBoostingClassifier = MPSBoostClassifier


# ------------------------------------------------------------------ single trees ---
_TREE_PARAMETERS = ("max_depth", "max_bins", "min_child_weight", "min_samples_leaf", "reg_lambda", "reg_alpha",
                    "max_delta_step", "monotonic_constraints", "interaction_constraints", "categorical_features",
                    "random_state", "device", "verbosity", "precision")


class _OneTree:
    """One tree with learning rate 1 (estimators/single_tree.py)."""

    _PARAMETER_NAMES = _TREE_PARAMETERS

    def __init__(self, max_depth: int = 6, max_bins: int = 256, min_child_weight: float = 1.0,
                 min_samples_leaf: int = 20, reg_lambda: float = 1.0, reg_alpha: float = 0.0,
                 max_delta_step: float = 0.0, monotonic_constraints: Any = None, interaction_constraints: Any = None,
                 categorical_features: Any = None, random_state: int | None = None, device: str = "hip",
                 verbosity: int = 1, precision: str = "fp64_fixed") -> None:
        super().__init__(n_estimators=1, learning_rate=1.0, max_depth=max_depth, max_bins=max_bins,
                         min_child_weight=min_child_weight, min_samples_leaf=min_samples_leaf,
                         reg_lambda=reg_lambda, reg_alpha=reg_alpha, max_delta_step=max_delta_step,
                         monotonic_constraints=monotonic_constraints, interaction_constraints=interaction_constraints,
                         categorical_features=categorical_features, random_state=random_state, device=device,
                         verbosity=verbosity, precision=precision)

    def fit(self, X: Any, y: Any, sample_weight: Any = None):
        self.n_estimators = 1
        self.learning_rate = 1.0
        return super().fit(X, y, sample_weight=sample_weight)

    def _validate_loaded_model(self, model: Any) -> None:
        if model.tree_count != 1:
            raise ValueError(f"{type(self).__name__} can only load one-tree models")


class DecisionTreeRegressor(_OneTree, MPSBoostRegressor):
    _fitted_error_message = "DecisionTreeRegressor is not fitted or loaded"


class ExtraTreeRegressor(DecisionTreeRegressor):
    _split_strategy = "random_threshold"
    _fitted_error_message = "ExtraTreeRegressor is not fitted or loaded"


class DecisionTreeClassifier(_OneTree, _BinaryClassifier):
    _fitted_error_message = "DecisionTreeClassifier is not fitted or loaded"


class ExtraTreeClassifier(DecisionTreeClassifier):
    _split_strategy = "random_threshold"
    _fitted_error_message = "ExtraTreeClassifier is not fitted or loaded"


# ----------------------------------------------------------------------- forests ---
_FOREST_PARAMETERS = ("n_estimators", "max_depth", "max_bins", "min_child_weight", "min_samples_leaf", "reg_lambda",
                      "reg_alpha", "max_delta_step", "monotonic_constraints", "interaction_constraints",
                      "categorical_features", "max_features", "sample_fraction", "bootstrap", "n_jobs",
                      "random_state", "device", "verbosity", "precision")


class _Forest:
    """Independent trees on sampled rows and features (estimators/forest.py:30-183,
    forest_sampling.py). The per-tree seeds, samples and feature subsets are the reference's;
    trees are trained one after another on the device's session."""

    _PARAMETER_NAMES = _FOREST_PARAMETERS
    _tree_type: type

    def __init__(self, n_estimators: int = 100, max_depth: int = 6, max_bins: int = 256,
                 min_child_weight: float = 1.0, min_samples_leaf: int = 20, reg_lambda: float = 1.0,
                 reg_alpha: float = 0.0, max_delta_step: float = 0.0, monotonic_constraints: Any = None,
                 interaction_constraints: Any = None, categorical_features: Any = None, max_features: float = 1.0,
                 sample_fraction: float = 1.0, bootstrap: bool = True, n_jobs: int = 1,
                 random_state: int | None = None, device: str = "hip", verbosity: int = 1,
                 precision: str = "fp64_fixed") -> None:
        super().__init__(n_estimators=n_estimators, learning_rate=1.0, max_depth=max_depth, max_bins=max_bins,
                         min_child_weight=min_child_weight, min_samples_leaf=min_samples_leaf,
                         reg_lambda=reg_lambda, reg_alpha=reg_alpha, max_delta_step=max_delta_step,
                         monotonic_constraints=monotonic_constraints, interaction_constraints=interaction_constraints,
                         categorical_features=categorical_features, random_state=random_state, device=device,
                         verbosity=verbosity, precision=precision)
        self.max_features = max_features
        self.sample_fraction = sample_fraction
        self.bootstrap = bootstrap
        self.n_jobs = n_jobs

    def _validate_forest_parameters(self) -> None:
        """The forest's own parameters (forest_sampling.py:118-150), then the per-tree
        parameters with the trees' learning rate of 1."""

        _run_checks(self, _FOREST_CHECKS)
        self.learning_rate = 1.0
        MPSBoostRegressor._validate_parameters(self)

    def _sample(self, labels: np.ndarray, random_state: int) -> tuple[np.ndarray, np.ndarray]:
        """A tree's rows and their labels."""

        rows = self._sample_rows(labels, random_state)
        return rows, self._take(labels, rows)

    def _take(self, labels: np.ndarray, rows: np.ndarray) -> np.ndarray:
        """labels[rows]. Labels that are all small whole numbers (class ids) are gathered from a
        one-byte copy, which stays in the host caches where the FP64 labels do not, and widened
        back: the same values."""

        codes = self.__dict__.get("_label_codes")
        if codes is not None and codes[0] is labels:
            return codes[1][rows].astype(np.float64)
        return labels[rows]

    def _sample_rows(self, labels: np.ndarray, random_state: int) -> np.ndarray:
        if self.bootstrap:
            return bootstrap_sample_indices(labels.shape[0], sample_fraction=float(self.sample_fraction),
                                            random_state=random_state)
        return sample_without_replacement_indices(labels.shape[0], sample_fraction=float(self.sample_fraction),
                                                  random_state=random_state)

    def _sample_features(self, n_features: int, random_state: int) -> np.ndarray:
        return subsample_feature_indices(n_features, feature_fraction=float(self.max_features),
                                         random_state=random_state)

    def _local_monotonic(self, features: np.ndarray) -> list[int] | None:
        if self.monotonic_constraints is None:
            return None
        raw = list(self.monotonic_constraints)
        return [raw[int(f)] for f in features]

    def _local_interactions(self, groups: list[list[int]], features: np.ndarray) -> list[list[int]] | None:
        if not groups:
            return None
        local = {int(f): i for i, f in enumerate(features)}
        out = [[local[f] for f in g if f in local] for g in groups]
        return [g for g in out if g]

    def fit(self, X: Any, y: Any, sample_weight: Any = None):
        if not self._fit_lock.acquire(blocking=False):
            raise RuntimeError("concurrent fit is not supported for one estimator")
        try:
            self._validate_forest_parameters()
            matrix = as_dense_matrix(X)
            labels = self._training_labels(y, matrix.shape[0])
            weights = as_sample_weight(sample_weight, matrix.shape[0])
            groups = self._normalized_interaction_constraints(matrix.shape[1])
            self._normalized_monotonic_constraints(matrix.shape[1])
            if labels.size and labels.min() >= 0.0 and labels.max() <= 255.0 and np.all(labels == np.floor(labels)):
                self._label_codes = (labels, labels.astype(np.uint8))
            self._decide_device(*matrix.shape)  # fails loudly without a device, before any upload
            resident = _native._DeviceMatrix(matrix)  # one upload; every tree bins its sample in place
            generator = np.random.default_rng(self.random_state)
            jobs = tuple((int(generator.integers(0, np.iinfo(np.int32).max)),
                          int(generator.integers(0, np.iinfo(np.int32).max))) for _ in range(self.n_estimators))

            unweighted = sample_weight is None

            def fit_one(seeds):
                row_seed, feature_seed = seeds
                rows, sample_labels = self._sample(labels, row_seed)
                features = self._sample_features(matrix.shape[1], feature_seed)
                tree = self._tree_type(max_depth=self.max_depth, max_bins=self.max_bins,
                                       min_child_weight=self.min_child_weight, min_samples_leaf=self.min_samples_leaf,
                                       reg_lambda=self.reg_lambda, reg_alpha=self.reg_alpha,
                                       max_delta_step=self.max_delta_step,
                                       monotonic_constraints=self._local_monotonic(features),
                                       interaction_constraints=self._local_interactions(groups, features),
                                       categorical_features=None, random_state=feature_seed, device=self.device,
                                       verbosity=self.verbosity, precision=self.precision)
                tree._fit_sampled(resident, rows, features, sample_labels, None if unweighted else weights[rows],
                                  labels_checked=True)
                return tree, features

            if self.n_jobs == 1:
                fitted = [fit_one(seeds) for seeds in jobs]
            else:
                # Each worker thread trains on its own device session (stream): trees of
                # uneven sizes overlap on the GPU.
                with ThreadPoolExecutor(max_workers=self.n_jobs) as pool:
                    fitted = list(pool.map(fit_one, jobs))
            estimators = [tree for tree, _ in fitted]
            subsets = [features for _, features in fitted]
            self.estimators_ = tuple(estimators)
            self.feature_subsets_ = tuple(subsets)
            self.n_features_in_ = matrix.shape[1]
            self.n_estimators_ = len(estimators)
            self.device_ = estimators[0].device_
            self.training_summary_ = {"n_estimators": len(estimators), "bootstrap": self.bootstrap,
                                      "sample_fraction": float(self.sample_fraction),
                                      "max_features": float(self.max_features), "n_jobs": self.n_jobs,
                                      "scheduling": "serial" if self.n_jobs == 1 else "thread_pool",
                                      "tree_devices": [e.device_ for e in estimators],
                                      "weighted": bool(sample_weight is not None), "precision": self.precision}
            self._finalize_fitted_metadata()
            return self
        except Exception:
            self._clear_fitted_state()
            raise
        finally:
            self.__dict__.pop("_label_codes", None)
            self._fit_lock.release()

    def _require_trees(self):
        if not hasattr(self, "estimators_"):
            raise NotFittedError(self._fitted_error_message)
        return self.estimators_

    def _checked(self, X: Any) -> np.ndarray:
        matrix = as_dense_matrix(X)
        if matrix.shape[1] != self.n_features_in_:
            raise ValueError("prediction feature count does not match training data")
        return matrix

    def feature_importance(self, kind: str = "gain") -> np.ndarray:
        """forest.py:185-199: per-tree importances mapped back to the original features."""

        self._require_trees()
        if kind not in {"gain", "split"}:
            raise ValueError("feature importance kind must be 'gain' or 'split'")
        values = np.zeros(self.n_features_in_, dtype=np.float64)
        for tree, features in zip(self.estimators_, self.feature_subsets_):
            values[features] += tree.feature_importance(kind=kind).astype(np.float64, copy=False)
        total = float(values.sum())
        if total > 0.0:
            values /= total
        return values.astype(np.float32, copy=False)

    def save_model(self, path: str) -> None:
        raise NotImplementedError("forest persistence is not provided by the hip estimators")


class RandomForestRegressor(_Forest, MPSBoostRegressor):
    _tree_type = DecisionTreeRegressor
    _fitted_error_message = "RandomForestRegressor is not fitted"

    def predict(self, X: Any) -> np.ndarray:
        """Mean of the trees' predictions (prediction.py:36-53)."""

        trees = self._require_trees()
        matrix = self._checked(X)
        out = np.zeros(matrix.shape[0], dtype=np.float64)
        for tree, features in zip(trees, self.feature_subsets_):
            out += tree.predict(matrix[:, features]).astype(np.float64, copy=False)
        out /= float(len(trees))
        return out.astype(np.float32)


class RandomForestClassifier(_Forest, _BinaryClassifier):
    _tree_type = DecisionTreeClassifier
    _fitted_error_message = "RandomForestClassifier is not fitted"

    def _sample(self, labels: np.ndarray, random_state: int) -> tuple[np.ndarray, np.ndarray]:
        """Retry until both classes are present (forest_public.py:49-56). The labels are 0 and
        1, so both are present exactly when their minimum and maximum differ."""

        for attempt in range(16):
            rows = self._sample_rows(labels, random_state + attempt)
            sample = self._take(labels, rows)
            if sample.min() != sample.max():
                return rows, sample
        return np.arange(labels.shape[0], dtype=np.int64), labels.copy()

    def predict_proba(self, X: Any) -> np.ndarray:
        """Mean of the trees' probabilities (prediction.py:56-73)."""

        trees = self._require_trees()
        matrix = self._checked(X)
        out = np.zeros((matrix.shape[0], 2), dtype=np.float64)
        for tree, features in zip(trees, self.feature_subsets_):
            out += tree.predict_proba(matrix[:, features]).astype(np.float64, copy=False)
        out /= float(len(trees))
        return out.astype(np.float32)


class ExtraTreesRegressor(RandomForestRegressor):
    _tree_type = ExtraTreeRegressor
    _fitted_error_message = "ExtraTreesRegressor is not fitted"


class ExtraTreesClassifier(RandomForestClassifier):
    _tree_type = ExtraTreeClassifier
    _fitted_error_message = "ExtraTreesClassifier is not fitted"


def fit_concurrently(jobs, n_jobs: int = 4) -> list:
    """Fit independent models: `jobs` holds (estimator, X, y) or (estimator, X, y, sample_weight)
    tuples. Each of the `n_jobs` worker threads trains on its own device session (a HIP stream
    with its own buffers), so fits of uneven sizes overlap on the GPU instead of queueing behind
    each other's host-side level work. Returns the fitted estimators in job order."""

    if isinstance(n_jobs, bool) or not isinstance(n_jobs, int) or n_jobs <= 0:
        raise ValueError("n_jobs must be a positive integer")

    def run(job):
        est, X, y = job[0], job[1], job[2]
        return est.fit(X, y, sample_weight=job[3] if len(job) > 3 else None)

    if n_jobs == 1:
        return [run(job) for job in jobs]
    with ThreadPoolExecutor(max_workers=n_jobs) as pool:
        return list(pool.map(run, jobs))


__all__ = [
    "DecisionTreeClassifier", "DecisionTreeRegressor", "ExtraTreeClassifier", "ExtraTreeRegressor",
    "ExtraTreesClassifier", "ExtraTreesRegressor", "GradientBoostingClassifier", "GradientBoostingRegressor",
    "MPSBoostClassifier", "MPSBoostRegressor", "NotFittedError", "RandomForestClassifier", "RandomForestRegressor",
    "as_dense_matrix", "as_labels", "as_sample_weight", "bootstrap_sample_indices", "fit_concurrently",
    "sample_without_replacement_indices", "subsample_feature_indices",
]
