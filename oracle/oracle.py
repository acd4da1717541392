"""ctypes driver for the CPU checkers under oracle/ (test infrastructure only).

Two libraries export the same C interface (``oracle_capi.h``):

* ``liboracle.so``: this repository's C++ restatement of the reference CPU trainer;
* ``_ref/libref.so``: the reference's own sources compiled unchanged (built only where
  the reference tree exists; it travels as a built file).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this module. The product (``mpsboost_amd``) never does.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REFERENCE_LIB = os.path.join(HERE, "_ref", "libref.so")
# The reference's trainers over the HIP backend (C ABI + include/mpsboost_hip_backend.hpp).
REFERENCE_HIP_LIB = os.path.join(HERE, "_ref", "libref_hip.so")

NODE_DTYPE = np.dtype(
    [
        ("feature_index", "<u4"),
        ("threshold_bin", "<u4"),
        ("left_child", "<u4"),
        ("right_child", "<u4"),
        ("leaf_value", "<f8"),
        ("gain", "<f8"),
        ("default_left", "u1"),
        ("flags", "u1"),
        ("pad", "u1", (6,)),
    ]
)
assert NODE_DTYPE.itemsize == 40

OBJECTIVES = {"squared_error": 0, "binary_logistic": 1, "quantile": 2, "poisson": 3, "tweedie": 4}


class DataError(ValueError):
    """Input data violates the binning contract (reference DataError)."""


class TrainingError(ValueError):
    """Training math or contract failure (reference TrainingError)."""


class _TreeParams(C.Structure):
    _fields_ = [
        ("max_depth", C.c_uint32),
        ("max_leaves", C.c_uint32),
        ("max_active_leaves", C.c_uint32),
        ("split_strategy", C.c_uint32),
        ("min_samples_leaf", C.c_uint64),
        ("min_child_weight", C.c_double),
        ("reg_lambda", C.c_double),
        ("reg_alpha", C.c_double),
        ("max_delta_step", C.c_double),
        ("gamma", C.c_double),
        ("min_gain_to_split", C.c_double),
        ("growth_strategy", C.c_uint32),
        ("random_seed", C.c_uint32),
        ("monotonic", C.POINTER(C.c_int8)),
        ("n_monotonic", C.c_uint32),
        ("n_groups", C.c_uint32),
        ("group_sizes", C.POINTER(C.c_uint32)),
        ("group_features", C.POINTER(C.c_uint32)),
    ]


class _TrainParams(C.Structure):
    _fields_ = [
        ("n_estimators", C.c_uint32),
        ("max_bins", C.c_uint32),
        ("objective", C.c_uint32),
        ("pad", C.c_uint32),
        ("learning_rate", C.c_double),
        ("objective_alpha", C.c_double),
        ("tweedie_variance_power", C.c_double),
        ("tree", _TreeParams),
    ]


def build(reference: bool = True) -> None:
    """Compile liboracle.so (and, when the reference tree exists, oracle/_ref/libref.so and
    oracle/_ref/libref_hip.so; the latter links the built product library)."""

    targets = ["oracle"] + (["ref", "ref_hip"] if reference else [])
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


def _ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(C.POINTER(ctype))


@dataclass
class TreeParams:
    """Tree growth parameters with the reference C++ defaults (include/mpsboost/tree.hpp:57-71)."""

    max_depth: int = 6
    max_leaves: int = 0
    max_active_leaves: int = 0
    min_samples_leaf: int = 1
    min_child_weight: float = 0.0
    reg_lambda: float = 1.0
    reg_alpha: float = 0.0
    max_delta_step: float = 0.0
    gamma: float = 0.0
    min_gain_to_split: float = 0.0
    split_strategy: str = "best_gain"
    growth_strategy: str = "level_wise"
    random_seed: int = 0
    monotonic_constraints: list = field(default_factory=list)
    interaction_constraints: list = field(default_factory=list)

    def to_c(self):
        keep = []
        p = _TreeParams()
        p.max_depth = self.max_depth
        p.max_leaves = self.max_leaves
        p.max_active_leaves = self.max_active_leaves
        p.split_strategy = {"best_gain": 0, "random_threshold": 1}[self.split_strategy]
        p.min_samples_leaf = self.min_samples_leaf
        p.min_child_weight = self.min_child_weight
        p.reg_lambda = self.reg_lambda
        p.reg_alpha = self.reg_alpha
        p.max_delta_step = self.max_delta_step
        p.gamma = self.gamma
        p.min_gain_to_split = self.min_gain_to_split
        p.growth_strategy = {"level_wise": 0, "leaf_wise": 1}[self.growth_strategy]
        p.random_seed = self.random_seed
        if len(self.monotonic_constraints):
            mono = np.ascontiguousarray(self.monotonic_constraints, dtype=np.int8)
            keep.append(mono)
            p.monotonic = _ptr(mono, C.c_int8)
            p.n_monotonic = mono.size
        if len(self.interaction_constraints):
            sizes = np.asarray([len(g) for g in self.interaction_constraints], dtype=np.uint32)
            flat = np.asarray([f for g in self.interaction_constraints for f in g], dtype=np.uint32)
            keep += [sizes, flat]
            p.n_groups = sizes.size
            p.group_sizes = _ptr(sizes, C.c_uint32)
            p.group_features = _ptr(flat, C.c_uint32)
        return p, keep


class Dataset:
    """Binned dataset handle (reference BinnedDataset view: bin 0 includes missing)."""

    def __init__(self, checker: "Checker", handle: int):
        self._checker = checker
        self._handle = C.c_void_p(handle)
        lib = checker.lib
        rows, feats, mb, bits, nb = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        lib.orc_dataset_info(self._handle, C.byref(rows), C.byref(feats), C.byref(mb), C.byref(bits), C.byref(nb))
        self.rows, self.features, self.max_bins = rows.value, feats.value, mb.value
        self.bin_width = bits.value
        self.bin_counts = np.zeros(self.features, dtype=np.uint32)
        self.boundary_offsets = np.zeros(self.features, dtype=np.uint64)
        self.missing_counts = np.zeros(self.features, dtype=np.uint64)
        flat = np.zeros(max(nb.value, 1), dtype=np.float32)
        lib.orc_dataset_meta(
            self._handle,
            _ptr(self.bin_counts, C.c_uint32),
            _ptr(self.boundary_offsets, C.c_uint64),
            _ptr(self.missing_counts, C.c_uint64),
            _ptr(flat, C.c_float),
        )
        self.boundary_values = flat[: nb.value]
        self._bins = None
        self._missing = None

    def __del__(self):
        try:
            self._checker.lib.orc_dataset_free(self._handle)
        except Exception:
            pass

    @property
    def handle(self):
        return self._handle

    @property
    def boundaries(self):
        out = []
        for f in range(self.features):
            lo = int(self.boundary_offsets[f])
            out.append(self.boundary_values[lo : lo + int(self.bin_counts[f]) - 1].tolist())
        return out

    def _load_bins(self):
        if self._bins is None:
            bins = np.zeros((self.features, self.rows), dtype=np.uint16)
            miss = np.zeros((self.features, self.rows), dtype=np.uint8)
            self._checker.lib.orc_dataset_bins(self._handle, _ptr(bins, C.c_uint16), _ptr(miss, C.c_uint8))
            self._bins, self._missing = bins, miss.astype(bool)

    @property
    def bins(self) -> np.ndarray:
        """Feature-major [features, rows] bins."""
        self._load_bins()
        return self._bins

    @property
    def missing(self) -> np.ndarray:
        self._load_bins()
        return self._missing


class Model:
    def __init__(self, checker: "Checker", handle: int):
        self._checker = checker
        self._handle = C.c_void_p(handle)
        trees, classes, lr = C.c_uint32(), C.c_uint32(), C.c_double()
        base = np.zeros(4096, dtype=np.float64)
        checker.lib.orc_model_info(self._handle, C.byref(trees), C.byref(classes), C.byref(lr), _ptr(base, C.c_double))
        self.tree_count, self.class_count, self.learning_rate = trees.value, classes.value, lr.value
        self.base_scores = base[: self.class_count].copy()

    def __del__(self):
        try:
            self._checker.lib.orc_model_free(self._handle)
        except Exception:
            pass

    def tree(self, index: int) -> np.ndarray:
        cap = 1 << 20
        nodes = np.zeros(cap, dtype=NODE_DTYPE)
        n = C.c_uint32()
        rc = self._checker.lib.orc_model_tree(self._handle, index, nodes.ctypes.data_as(C.c_void_p), cap, C.byref(n))
        if rc:
            raise IndexError(index)
        return nodes[: n.value].copy()

    @property
    def trees(self):
        return [self.tree(i) for i in range(self.tree_count)]

    def serialize(self) -> bytes:
        """The reference's SerializeModel bytes (compiled reference only)."""

        size = C.c_uint64()
        self._checker._call("orc_model_serialize", self._handle, None, C.c_uint64(0), C.byref(size))
        out = (C.c_uint8 * size.value)()
        self._checker._call("orc_model_serialize", self._handle, out, C.c_uint64(size.value), C.byref(size))
        return bytes(out)

    def predict(self, dataset: Dataset) -> np.ndarray:
        out = np.zeros(dataset.rows * (self.class_count if self.class_count > 1 else 1), dtype=np.float64)
        self._checker._call("orc_model_predict", self._handle, dataset.handle, _ptr(out, C.c_double))
        return out if self.class_count == 1 else out.reshape(dataset.rows, self.class_count)


class Checker:
    """One CPU checker library (the restatement or the compiled reference)."""

    def __init__(self, path: str = ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = C.CDLL(path)
        self.lib.orc_dataset_free.restype = None
        self.lib.orc_model_free.restype = None

    # -- plumbing
    def set_threads(self, threads: int) -> None:
        """Threads for the oracle's per-feature loops; results are bit-identical for any count
        (liboracle.so only; the compiled reference has no threads)."""

        fn = getattr(self.lib, "orc_set_threads", None)
        if fn is None:
            raise AttributeError("this checker has no threads setting")
        fn(C.c_uint32(int(threads)))

    def _call(self, name, *args):
        err = C.create_string_buffer(512)
        rc = getattr(self.lib, name)(*args, err, C.c_uint64(512))
        if rc == 1:
            raise DataError(err.value.decode())
        if rc == 2:
            raise TrainingError(err.value.decode())
        if rc:
            raise ValueError(err.value.decode())

    @staticmethod
    def _matrix(matrix):
        m = np.asarray(matrix)
        if m.dtype not in (np.float32, np.float64):
            raise TypeError("Input dtype must be float32 or float64")
        if m.ndim != 2:
            raise ValueError("Input matrix must be two-dimensional")
        return m, (1 if m.dtype == np.float32 else 2)

    # -- binning
    def quantize(self, matrix, max_bins: int) -> Dataset:
        m, scalar = self._matrix(matrix)
        out = C.c_void_p()
        self._call(
            "orc_quantize", m.ctypes.data_as(C.c_void_p), C.c_uint64(m.shape[0]), C.c_uint32(m.shape[1]),
            C.c_uint64(m.strides[0]), C.c_uint64(m.strides[1]), C.c_int(scalar), C.c_uint32(max_bins), C.byref(out),
        )
        return Dataset(self, out.value)

    def transform(self, matrix, schema: Dataset) -> Dataset:
        m, scalar = self._matrix(matrix)
        out = C.c_void_p()
        self._call(
            "orc_transform", m.ctypes.data_as(C.c_void_p), C.c_uint64(m.shape[0]), C.c_uint32(m.shape[1]),
            C.c_uint64(m.strides[0]), C.c_uint64(m.strides[1]), C.c_int(scalar), schema.handle, C.byref(out),
        )
        return Dataset(self, out.value)

    # -- objective math
    def squared_error(self, labels, predictions) -> np.ndarray:
        y = np.ascontiguousarray(labels, dtype=np.float64)
        p = np.ascontiguousarray(predictions, dtype=np.float64)
        if y.size != p.size:
            raise TrainingError("Label and prediction lengths do not match")
        out = np.zeros((y.size, 2), dtype=np.float64)
        self._call("orc_squared_error", _ptr(y, C.c_double), _ptr(p, C.c_double), C.c_uint64(y.size), _ptr(out, C.c_double))
        return out

    def gradients(self, objective, labels, predictions, alpha=0.5, class_count=0, target_class=0):
        y = np.ascontiguousarray(labels, dtype=np.float64)
        p = np.ascontiguousarray(predictions, dtype=np.float64)
        code = 5 if objective == "softmax" else OBJECTIVES[objective]
        out = np.zeros((y.size, 2), dtype=np.float64)
        self._call(
            "orc_gradients", C.c_uint32(code), _ptr(y, C.c_double), _ptr(p, C.c_double), C.c_uint64(y.size),
            C.c_double(alpha), C.c_uint32(class_count), C.c_uint32(target_class), _ptr(out, C.c_double),
        )
        return out

    def node_score(self, g, h, reg_lambda, reg_alpha=0.0):
        out = C.c_double()
        self._call("orc_node_score", C.c_double(g), C.c_double(h), C.c_double(reg_lambda), C.c_double(reg_alpha), C.byref(out))
        return out.value

    def leaf_weight(self, g, h, reg_lambda, reg_alpha=0.0, max_delta_step=0.0):
        out = C.c_double()
        self._call(
            "orc_leaf_weight", C.c_double(g), C.c_double(h), C.c_double(reg_lambda), C.c_double(reg_alpha),
            C.c_double(max_delta_step), C.byref(out),
        )
        return out.value

    def split_gain(self, gl, hl, gr, hr, reg_lambda, gamma, reg_alpha=0.0):
        out = C.c_double()
        self._call(
            "orc_split_gain", C.c_double(gl), C.c_double(hl), C.c_double(gr), C.c_double(hr),
            C.c_double(reg_lambda), C.c_double(reg_alpha), C.c_double(gamma), C.byref(out),
        )
        return out.value

    # -- histograms and trees
    def histograms(self, dataset: Dataset, pairs, rows):
        """Reference-view histogram: returns (count, G, H) flat over all cells."""

        gh = np.ascontiguousarray(pairs, dtype=np.float64).reshape(-1, 2)
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        cells = int(dataset.bin_counts.sum())
        cnt = np.zeros(cells, dtype=np.uint64)
        g = np.zeros(cells, dtype=np.float64)
        h = np.zeros(cells, dtype=np.float64)
        self._call(
            "orc_histograms", dataset.handle, _ptr(gh, C.c_double), _ptr(r, C.c_uint64), C.c_uint64(r.size),
            _ptr(cnt, C.c_uint64), _ptr(g, C.c_double), _ptr(h, C.c_double),
        )
        return cnt, g, h

    def train_tree(self, dataset: Dataset, pairs, params: TreeParams | None = None, **kw) -> np.ndarray:
        params = params or TreeParams(**kw)
        gh = np.ascontiguousarray(pairs, dtype=np.float64).reshape(-1, 2)
        cp, keep = params.to_c()
        cap = 1 << 20
        nodes = np.zeros(cap, dtype=NODE_DTYPE)
        n = C.c_uint32()
        self._call(
            "orc_train_tree", dataset.handle, _ptr(gh, C.c_double), C.byref(cp), nodes.ctypes.data_as(C.c_void_p),
            C.c_uint32(cap), C.byref(n),
        )
        del keep
        return nodes[: n.value].copy()

    def tree_predict(self, dataset: Dataset, nodes: np.ndarray, tree_features: int | None = None) -> np.ndarray:
        """Predict one tree; ``tree_features`` is the feature count it was trained on."""

        nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        out = np.zeros(dataset.rows, dtype=np.float64)
        feats = dataset.features if tree_features is None else tree_features
        self._call("orc_tree_predict", dataset.handle, nodes.ctypes.data_as(C.c_void_p), C.c_uint32(nodes.size),
                   C.c_uint32(feats), _ptr(out, C.c_double))
        return out

    def _train_params(self, n_estimators, learning_rate, max_bins, objective, objective_alpha,
                      tweedie_variance_power, tree):
        cp, keep = tree.to_c()
        tp = _TrainParams()
        tp.n_estimators = n_estimators
        tp.max_bins = max_bins
        tp.objective = OBJECTIVES[objective]
        tp.learning_rate = learning_rate
        tp.objective_alpha = objective_alpha
        tp.tweedie_variance_power = tweedie_variance_power
        tp.tree = cp
        return tp, keep

    def train_regressor(self, dataset: Dataset, labels, weights=None, *, n_estimators=100, learning_rate=0.1,
                        objective="squared_error", objective_alpha=0.5, tweedie_variance_power=1.5,
                        tree: TreeParams | None = None, **tree_kw) -> Model:
        y = np.ascontiguousarray(labels, dtype=np.float64)
        w = np.ones_like(y) if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        tp, keep = self._train_params(n_estimators, learning_rate, dataset.max_bins, objective, objective_alpha,
                                      tweedie_variance_power, tree or TreeParams(**tree_kw))
        out = C.c_void_p()
        self._call("orc_train_regressor", dataset.handle, _ptr(y, C.c_double), _ptr(w, C.c_double), C.byref(tp), C.byref(out))
        del keep
        return Model(self, out.value)

    def deserialize_model(self, data: bytes) -> Model:
        """The reference's DeserializeModel / DeserializeMulticlassModel (compiled reference only)."""

        buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
        out = C.c_void_p()
        self._call("orc_model_deserialize", buf, C.c_uint64(len(data)), C.byref(out))
        return Model(self, out.value)

    def train_multiclass(self, dataset: Dataset, labels, class_count: int, weights=None, *, n_estimators=100,
                         learning_rate=0.1, tree: TreeParams | None = None, **tree_kw) -> Model:
        y = np.ascontiguousarray(labels, dtype=np.float64)
        w = np.ones_like(y) if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        tp, keep = self._train_params(n_estimators, learning_rate, dataset.max_bins, "squared_error", 0.5, 1.5,
                                      tree or TreeParams(**tree_kw))
        out = C.c_void_p()
        self._call("orc_train_multiclass", dataset.handle, _ptr(y, C.c_double), _ptr(w, C.c_double), C.byref(tp),
                   C.c_uint32(class_count), C.byref(out))
        del keep
        return Model(self, out.value)


_cache: dict[str, Checker] = {}


def oracle() -> Checker:
    """The repository's own CPU restatement."""

    if ORACLE_LIB not in _cache:
        _cache[ORACLE_LIB] = Checker(ORACLE_LIB)
    return _cache[ORACLE_LIB]


def reference() -> Checker | None:
    """The compiled reference, or None where it was never built."""

    if not os.path.exists(REFERENCE_LIB):
        return None
    if REFERENCE_LIB not in _cache:
        _cache[REFERENCE_LIB] = Checker(REFERENCE_LIB)
    return _cache[REFERENCE_LIB]


def reference_hip() -> Checker | None:
    """The reference's own trainers driving the HIP backend through its C++ interfaces
    (needs a GPU to train), or None where it was never built."""

    if not os.path.exists(REFERENCE_HIP_LIB):
        return None
    if REFERENCE_HIP_LIB not in _cache:
        _cache[REFERENCE_HIP_LIB] = Checker(REFERENCE_HIP_LIB)
    return _cache[REFERENCE_HIP_LIB]
