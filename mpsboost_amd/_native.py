"""ctypes binding of libmpsboost_hip.so, mirroring the reference's ``_native`` module.

The names follow the reference's pybind11 surface (src/python/binding_*.cpp): ``_quantize_dense``,
``_TrainingParameters``, ``_train_regressor_*``, ``_backend_info`` and a ``_HipBackend`` test
class with the same hook methods as ``_MpsBackend`` (gradients, histograms,
baseline_histograms, split_candidates, partition_rows, last_timing). Everything crosses the
C ABI declared in include/mpsboost_hip.h; there is no CPU fallback: without the built library
or without a gfx950 device the calls raise.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
from typing import Any

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "lib", "libmpsboost_hip.so")

PRECISIONS = {"fp64": 0, "fp32": 1, "int16": 2, "int8": 3, "fp64_fixed": 4}
OBJECTIVES = {"squared_error": 0}

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


class BackendError(RuntimeError):
    """Device/runtime failure with stage context (reference BackendError)."""


class DataError(ValueError):
    """Input data violates the binning contract (reference DataError)."""


class TrainingError(ValueError):
    """Math or contract failure during training (reference TrainingError)."""


class _BackendInfo(C.Structure):
    _fields_ = [
        ("available", C.c_int32),
        ("compute_units", C.c_uint32),
        ("total_memory", C.c_uint64),
        ("device_name", C.c_char * 96),
        ("arch", C.c_char * 32),
    ]


class _Timing(C.Structure):
    _fields_ = [
        ("gradient_seconds", C.c_double),
        ("histogram_encode_seconds", C.c_double),
        ("histogram_command_seconds", C.c_double),
        ("hot_path_encode_seconds", C.c_double),
        ("hot_path_command_seconds", C.c_double),
        ("pooled_buffer_reuse_count", C.c_uint64),
        ("pooled_buffer_allocation_count", C.c_uint64),
    ]


class _DatasetInfo(C.Structure):
    _fields_ = [
        ("rows", C.c_uint64),
        ("features", C.c_uint32),
        ("max_bins", C.c_uint32),
        ("storage_bits", C.c_uint32),
        ("code_bytes", C.c_uint32),
        ("row_stride", C.c_uint32),
        ("cells", C.c_uint32),
        ("boundary_total", C.c_uint64),
    ]


class _SplitParams(C.Structure):
    _fields_ = [
        ("min_samples_leaf", C.c_uint64),
        ("min_child_weight", C.c_double),
        ("reg_lambda", C.c_double),
        ("reg_alpha", C.c_double),
        ("gamma", C.c_double),
        ("min_gain_to_split", C.c_double),
    ]


class _SplitCandidate(C.Structure):
    _fields_ = [
        ("valid", C.c_uint32),
        ("feature", C.c_uint32),
        ("threshold_bin", C.c_uint32),
        ("default_left", C.c_uint32),
        ("left_count", C.c_uint64),
        ("right_count", C.c_uint64),
        ("left_gradient_sum", C.c_double),
        ("left_hessian_sum", C.c_double),
        ("right_gradient_sum", C.c_double),
        ("right_hessian_sum", C.c_double),
        ("gain", C.c_double),
    ]


class _TrainParamsC(C.Structure):
    _fields_ = [
        ("n_estimators", C.c_uint32),
        ("max_bins", C.c_uint32),
        ("objective", C.c_uint32),
        ("precision", C.c_uint32),
        ("learning_rate", C.c_double),
        ("objective_alpha", C.c_double),
        ("tweedie_variance_power", C.c_double),
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
        ("rounding_seed", C.c_uint64),
        ("leaf_solve", C.c_uint32),
        ("near_ties", C.c_uint32),
        ("class_rounds", C.c_uint32),
    ]


class _ModelInfo(C.Structure):
    _fields_ = [
        ("tree_count", C.c_uint32),
        ("class_count", C.c_uint32),
        ("features", C.c_uint32),
        ("precision", C.c_uint32),
        ("learning_rate", C.c_double),
        ("base_scores", C.c_double * 256),
        ("objective", C.c_uint32),
        ("max_bins", C.c_uint32),
        ("objective_alpha", C.c_double),
        ("tweedie_variance_power", C.c_double),
        ("class_labels", C.c_double * 256),
    ]


OBJECTIVE_NAMES = ("squared_error", "binary_logistic", "quantile", "poisson", "tweedie")


class _FitStats(C.Structure):
    _fields_ = [
        ("trees", C.c_uint64),
        ("histogram_rows", C.c_uint64),
        ("histogram_bytes", C.c_uint64),
        ("histogram_launches", C.c_uint64),
        ("histogram_seconds", C.c_double),
        ("scan_seconds", C.c_double),
        ("partition_seconds", C.c_double),
        ("gradient_seconds", C.c_double),
        ("reduce_seconds", C.c_double),
        ("train_seconds", C.c_double),
        ("levels", C.c_uint64),
        ("timed_trees", C.c_uint64),
        ("histogram_timed_bytes", C.c_uint64),
        ("histogram_timed_launches", C.c_uint64),
        ("device_planned_trees", C.c_uint64),
        ("renewal_seconds", C.c_double),
        ("renewed_trees", C.c_uint64),
        ("ambiguous_nodes", C.c_uint64),
        ("resolved_nodes", C.c_uint64),
        ("resolved_features", C.c_uint64),
        ("setup_seconds", C.c_double),
        ("finish_seconds", C.c_double),
        ("base_score_seconds", C.c_double),
        ("root_gradient_only_builds", C.c_uint64),
    ]


_lib_lock = threading.Lock()
_lib_handle = None


def _lib():
    """Load the built library; raise BackendError (never fall back) when it is missing."""

    global _lib_handle
    with _lib_lock:
        if _lib_handle is None:
            if not os.path.exists(LIBRARY):
                raise BackendError(
                    f"{LIBRARY} is not built: run `python -m mpsboost_amd.build` (hipcc, gfx950)"
                )
            lib = C.CDLL(LIBRARY)
            lib.mpsb_last_error.restype = C.c_char_p
            for name in ("mpsb_session_destroy", "mpsb_dataset_free", "mpsb_model_free"):
                getattr(lib, name).restype = None
            _lib_handle = lib
        return _lib_handle


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = (_lib().mpsb_last_error() or b"").decode()
    if rc == 1:
        raise DataError(msg)
    if rc == 2:
        raise TrainingError(msg)
    if rc == 3:
        raise ValueError(msg)
    raise BackendError(msg)


def _p(array: np.ndarray, ctype):
    return array.ctypes.data_as(C.POINTER(ctype))


def _matrix(matrix) -> tuple[np.ndarray, int]:
    m = np.asarray(matrix)
    if m.dtype not in (np.float32, np.float64):
        raise TypeError("Input dtype must be native float32 or float64")
    if m.ndim != 2:
        raise ValueError("Input matrix must be a two-dimensional buffer")
    if m.shape[0] <= 0 or m.shape[1] <= 0:
        raise ValueError("Input matrix must contain at least one row and one feature")
    if m.strides[0] <= 0 or m.strides[1] <= 0:
        raise ValueError("Zero or negative strides are not supported")
    return m, (1 if m.dtype == np.float32 else 2)


# ------------------------------------------------------------------ device query ----
def _backend_info() -> dict[str, Any]:
    """QueryBackendInfo: availability is a value, not an exception."""

    info = _BackendInfo()
    _check(_lib().mpsb_query_backend_info(C.byref(info)))
    return {
        "available": bool(info.available),
        "device_name": info.device_name.decode() if info.available else None,
        "arch": info.arch.decode(),
        "compute_units": info.compute_units,
        "total_memory": info.total_memory,
    }


# -------------------------------------------------------------------- datasets ------
class _BinnedDataset:
    """Binned dataset (host copy of the schema and the row-major device codes)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        info = _DatasetInfo()
        _check(_lib().mpsb_dataset_info_get(self._h, C.byref(info)))
        self.rows = info.rows
        self.features = info.features
        self.max_bins = info.max_bins
        self.bin_width = info.storage_bits
        self.code_bytes = info.code_bytes
        self.row_stride = info.row_stride
        self.cells = info.cells
        self._boundary_total = info.boundary_total
        self.bin_counts = np.zeros(self.features, dtype=np.uint32)
        self.boundary_offsets = np.zeros(self.features, dtype=np.uint64)
        self.missing_counts = np.zeros(self.features, dtype=np.uint64)
        self.boundary_values = np.zeros(max(1, self._boundary_total), dtype=np.float32)
        _check(
            _lib().mpsb_dataset_meta(
                self._h, _p(self.bin_counts, C.c_uint32), _p(self.boundary_offsets, C.c_uint64),
                _p(self.missing_counts, C.c_uint64), _p(self.boundary_values, C.c_float),
            )
        )
        self.boundary_values = self.boundary_values[: self._boundary_total]
        self._bins = None

    def __del__(self):
        try:
            _lib().mpsb_dataset_free(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def boundaries(self) -> list[list[float]]:
        out = []
        for f in range(self.features):
            lo = int(self.boundary_offsets[f])
            out.append(self.boundary_values[lo : lo + int(self.bin_counts[f]) - 1].tolist())
        return out

    def _load(self):
        if self._bins is None:
            bins = np.zeros((self.features, self.rows), dtype=np.uint16)
            miss = np.zeros((self.features, self.rows), dtype=np.uint8)
            _check(_lib().mpsb_dataset_bins(self._h, _p(bins, C.c_uint16), _p(miss, C.c_uint8)))
            self._bins, self._missing = bins, miss.astype(bool)

    def device_codes(self) -> np.ndarray:
        """The codes in the device layout: uint8 [rows, row_stride * code_bytes], row-major
        (mpsb_dataset_codes)."""

        size = C.c_uint64(0)
        _check(_lib().mpsb_dataset_codes(self._h, None, C.c_uint64(0), C.byref(size)))
        out = np.empty(size.value, dtype=np.uint8)
        _check(_lib().mpsb_dataset_codes(self._h, out.ctypes.data_as(C.c_void_p), size, C.byref(size)))
        return out.reshape(self.rows, -1)

    @property
    def bins(self) -> np.ndarray:
        """Reference (feature-major) bins [features, rows]; missing rows read 0."""

        self._load()
        return self._bins

    @property
    def missing(self) -> np.ndarray:
        self._load()
        return self._missing


def _build_histograms_noalloc(dataset: _BinnedDataset, codes: int, rows: int, n_rows: int, pairs: int,
                              precision: str, count: int, gradient: int, hessian: int, stream: int = 0) -> None:
    """mpsb_build_histograms_noalloc on device addresses (integers): codes in the device layout,
    uint64 rows, FP64 pairs [2 * dataset rows], outputs uint64 / FP64 / FP64 [reference cells];
    stream is a hipStream_t (0: the null stream)."""

    _check(_lib().mpsb_build_histograms_noalloc(
        dataset.handle, C.c_void_p(codes), C.c_void_p(rows), C.c_uint64(n_rows), C.c_void_p(pairs),
        C.c_int32(PRECISIONS[precision]), C.c_void_p(count), C.c_void_p(gradient), C.c_void_p(hessian),
        C.c_void_p(stream)))


def _quantize_dense(matrix, max_bins: int, threads: int = 0) -> _BinnedDataset:
    """QuantizeDense (src/core/binned_dataset.cpp:126-188) on the host, multi-threaded."""

    m, scalar = _matrix(matrix)
    out = C.c_void_p()
    _check(
        _lib().mpsb_quantize_dense(
            m.ctypes.data_as(C.c_void_p), C.c_uint64(m.shape[0]), C.c_uint32(m.shape[1]), C.c_uint64(m.strides[0]),
            C.c_uint64(m.strides[1]), C.c_int32(scalar), C.c_uint32(max_bins), C.c_uint32(threads), C.byref(out),
        )
    )
    return _BinnedDataset(out)


def _transform_dense(matrix, schema: _BinnedDataset, threads: int = 0) -> _BinnedDataset:
    """TransformDense (src/core/binned_dataset.cpp:190-238): frozen boundaries."""

    m, scalar = _matrix(matrix)
    out = C.c_void_p()
    _check(
        _lib().mpsb_transform_dense(
            m.ctypes.data_as(C.c_void_p), C.c_uint64(m.shape[0]), C.c_uint32(m.shape[1]), C.c_uint64(m.strides[0]),
            C.c_uint64(m.strides[1]), C.c_int32(scalar), schema.handle, C.c_uint32(threads), C.byref(out),
        )
    )
    return _BinnedDataset(out)


def _quantize_dense_device(session: "_Session", matrix, max_bins: int) -> _BinnedDataset:
    """QuantizeDense (src/core/binned_dataset.cpp:126-188) on the session's GPU; the codes stay
    resident there for the next fit."""

    m, scalar = _matrix(matrix)
    out = C.c_void_p()
    _check(
        _lib().mpsb_quantize_dense_device(
            session.handle, m.ctypes.data_as(C.c_void_p), C.c_uint64(m.shape[0]), C.c_uint32(m.shape[1]),
            C.c_uint64(m.strides[0]), C.c_uint64(m.strides[1]), C.c_int32(scalar), C.c_uint32(max_bins), C.byref(out),
        )
    )
    return _BinnedDataset(out)


def _transform_dense_device(session: "_Session", matrix, schema: _BinnedDataset) -> _BinnedDataset:
    """TransformDense (src/core/binned_dataset.cpp:190-238) on the session's GPU."""

    m, scalar = _matrix(matrix)
    out = C.c_void_p()
    _check(
        _lib().mpsb_transform_dense_device(
            session.handle, m.ctypes.data_as(C.c_void_p), C.c_uint64(m.shape[0]), C.c_uint32(m.shape[1]),
            C.c_uint64(m.strides[0]), C.c_uint64(m.strides[1]), C.c_int32(scalar), schema.handle, C.byref(out),
        )
    )
    return _BinnedDataset(out)


class _DeviceMatrix:
    """A dense matrix copied once to device memory, for sampled binning of many trees."""

    def __init__(self, matrix, device: int = 0):
        m, scalar = _matrix(matrix)
        self._h = C.c_void_p()
        self.rows, self.features = m.shape
        _check(_lib().mpsb_matrix_upload(C.c_int32(device), m.ctypes.data_as(C.c_void_p), C.c_uint64(m.shape[0]),
                                         C.c_uint32(m.shape[1]), C.c_uint64(m.strides[0]), C.c_uint64(m.strides[1]),
                                         C.c_int32(scalar), C.byref(self._h)))

    def __del__(self):
        try:
            if self._h:
                _lib().mpsb_matrix_free(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h


def _quantize_matrix_sampled(session: "_Session", matrix: _DeviceMatrix, rows, features,
                             max_bins: int) -> _BinnedDataset:
    """QuantizeDense of matrix[rows][:, features], read in place on the device."""

    rows = np.asarray(rows)
    if rows.dtype == np.int64 and rows.flags.c_contiguous and (rows.size == 0 or rows.min() >= 0):
        r = rows.view(np.uint64)  # the same bits for non-negative indices: no copy
    else:
        r = np.ascontiguousarray(rows, dtype=np.uint64)
    f = np.ascontiguousarray(features, dtype=np.uint32)
    out = C.c_void_p()
    _check(_lib().mpsb_quantize_matrix_sampled(session.handle, matrix.handle, _p(r, C.c_uint64), C.c_uint64(r.size),
                                               _p(f, C.c_uint32), C.c_uint32(f.size), C.c_uint32(max_bins),
                                               C.byref(out)))
    return _BinnedDataset(out)


def _dataset_from_bins(bins: np.ndarray, bin_counts, max_bins: int, missing: np.ndarray | None = None,
                       boundaries=None) -> _BinnedDataset:
    """Adopt feature-major [features, rows] bins of an existing binned dataset."""

    b = np.ascontiguousarray(bins)
    if b.dtype not in (np.uint8, np.uint16):
        raise TypeError("bins must be uint8 or uint16")
    counts = np.ascontiguousarray(bin_counts, dtype=np.uint32)
    miss = None if missing is None else np.ascontiguousarray(missing, dtype=np.uint8)
    bnd = None if boundaries is None else np.ascontiguousarray(boundaries, dtype=np.float32)
    out = C.c_void_p()
    _check(
        _lib().mpsb_dataset_from_bins(
            C.c_uint64(b.shape[1]), C.c_uint32(b.shape[0]), C.c_uint32(max_bins), _p(counts, C.c_uint32),
            b.ctypes.data_as(C.c_void_p), C.c_int32(b.itemsize), None if miss is None else _p(miss, C.c_uint8),
            None if bnd is None else _p(bnd, C.c_float), C.byref(out),
        )
    )
    return _BinnedDataset(out)


# -------------------------------------------------------------- training params -----
LEAF_SOLVE = {"auto": 0, "histogram": 1, "renewed": 2}
NEAR_TIES = {"auto": 0, "resolve": 1, "keep": 2}
CLASS_ROUNDS = {"class_order": 0, "batched": 1}


class _TrainingParameters:
    """MakeTrainingParameters (src/python/binding_dataset.cpp:23-103) plus the device's
    histogram ``precision`` ("fp64", "fp32", "int16", "int8"), ``rounding_seed`` and
    ``leaf_solve`` ("auto", "histogram" or "renewed": leaf values from exact FP64 gradient sums
    once a tree is grown; "auto" renews in the fp32, int16 and int8 forms) and ``near_ties``
    ("auto", "resolve" or "keep": int16 / int8 splits that rounding noise could have swapped
    are scanned again exactly; "auto" resolves in int16 and keeps in int8) and, for softmax,
    ``class_rounds`` ("class_order": the reference's order, each class tree after the earlier
    class trees of its round; "batched": every class's gradients from the round's starting
    margins, the round's class trees grown together, one launch per level for all of them)."""

    def __init__(self, n_estimators, learning_rate, max_bins, max_depth, min_samples_leaf, min_child_weight,
                 reg_lambda, reg_alpha=0.0, max_delta_step=0.0, gamma=0.0, min_gain_to_split=0.0,
                 objective="squared_error", split_strategy="best_gain", growth_strategy="level_wise",
                 max_leaves=0, max_active_leaves=0, random_seed=0, monotonic_constraints=(),
                 interaction_constraints=(), objective_alpha=0.5, tweedie_variance_power=1.5,
                 precision="fp64", rounding_seed=0, leaf_solve="auto", near_ties="auto",
                 class_rounds="class_order"):
        if class_rounds not in CLASS_ROUNDS:
            raise ValueError(f"class_rounds must be one of {', '.join(CLASS_ROUNDS)}")
        if near_ties not in NEAR_TIES:
            raise ValueError(f"near_ties must be one of {', '.join(NEAR_TIES)}")
        if leaf_solve not in LEAF_SOLVE:
            raise ValueError(f"leaf_solve must be one of {', '.join(LEAF_SOLVE)}")
        if objective not in ("squared_error", "binary_logistic", "quantile", "poisson", "tweedie"):
            raise ValueError("unknown training objective")
        if split_strategy not in ("best_gain", "random_threshold"):
            raise ValueError("unknown split strategy")
        if growth_strategy not in ("level_wise", "leaf_wise"):
            raise ValueError("unknown growth strategy")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.n_estimators = int(n_estimators)
        self.learning_rate = float(learning_rate)
        self.max_bins = int(max_bins)
        self.max_depth = int(max_depth)
        self.min_samples_leaf = int(min_samples_leaf)
        self.min_child_weight = float(min_child_weight)
        self.reg_lambda = float(reg_lambda)
        self.reg_alpha = float(reg_alpha)
        self.max_delta_step = float(max_delta_step)
        self.gamma = float(gamma)
        self.min_gain_to_split = float(min_gain_to_split)
        self.objective = objective
        self.split_strategy = split_strategy
        self.growth_strategy = growth_strategy
        self.max_leaves = int(max_leaves)
        self.max_active_leaves = int(max_active_leaves)
        self.random_seed = int(random_seed)
        self.monotonic_constraints = list(monotonic_constraints)
        self.interaction_constraints = [list(g) for g in interaction_constraints]
        self.objective_alpha = float(objective_alpha)
        self.tweedie_variance_power = float(tweedie_variance_power)
        self.precision = precision
        self.rounding_seed = int(rounding_seed)
        self.leaf_solve = leaf_solve
        self.near_ties = near_ties
        self.class_rounds = class_rounds

    def to_c(self):
        keep = []
        p = _TrainParamsC()
        p.n_estimators = self.n_estimators
        p.max_bins = self.max_bins
        p.objective = {"squared_error": 0, "binary_logistic": 1, "quantile": 2, "poisson": 3, "tweedie": 4}[self.objective]
        p.precision = PRECISIONS[self.precision]
        p.learning_rate = self.learning_rate
        p.objective_alpha = self.objective_alpha
        p.tweedie_variance_power = self.tweedie_variance_power
        p.max_depth = self.max_depth
        p.max_leaves = self.max_leaves
        p.max_active_leaves = self.max_active_leaves
        p.split_strategy = 1 if self.split_strategy == "random_threshold" else 0
        if self.min_samples_leaf < 0:
            raise TrainingError("min_samples_leaf must be at least 1")
        p.min_samples_leaf = self.min_samples_leaf
        p.min_child_weight = self.min_child_weight
        p.reg_lambda = self.reg_lambda
        p.reg_alpha = self.reg_alpha
        p.max_delta_step = self.max_delta_step
        p.gamma = self.gamma
        p.min_gain_to_split = self.min_gain_to_split
        p.growth_strategy = 1 if self.growth_strategy == "leaf_wise" else 0
        p.random_seed = self.random_seed & 0xFFFFFFFF
        if self.monotonic_constraints:
            mono = np.ascontiguousarray(self.monotonic_constraints, dtype=np.int8)
            keep.append(mono)
            p.monotonic = _p(mono, C.c_int8)
            p.n_monotonic = mono.size
        if self.interaction_constraints:
            sizes = np.asarray([len(g) for g in self.interaction_constraints], dtype=np.uint32)
            flat = np.asarray([f for g in self.interaction_constraints for f in g], dtype=np.uint32)
            keep += [sizes, flat]
            p.n_groups = sizes.size
            p.group_sizes = _p(sizes, C.c_uint32)
            p.group_features = _p(flat, C.c_uint32)
        p.rounding_seed = self.rounding_seed & 0xFFFFFFFFFFFFFFFF
        p.leaf_solve = LEAF_SOLVE[self.leaf_solve]
        p.near_ties = NEAR_TIES[self.near_ties]
        p.class_rounds = CLASS_ROUNDS[self.class_rounds]
        return p, keep


# ------------------------------------------------------------------- sessions -------
class _Tuning(C.Structure):
    _fields_ = [
        ("lds_budget", C.c_uint32),
        ("target_blocks", C.c_uint32),
        ("min_slab_rows", C.c_uint32),
        ("block", C.c_uint32),
        ("words_per_thread", C.c_uint32),
        ("planner", C.c_uint32),
    ]


class _Session:
    """One device session: a HIP stream and its pooled buffers. Sessions are independent and
    may run concurrently from different threads; one session serves one thread at a time."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        self.device = device
        _check(_lib().mpsb_session_create(C.c_int32(device), C.byref(self._h)))

    def _get_tuning(self) -> dict[str, int]:
        t = _Tuning()
        _check(_lib().mpsb_session_get_tuning(self._h, C.byref(t)))
        return {name: int(getattr(t, name)) for name, _ in _Tuning._fields_}

    def _set_tuning(self, values: dict[str, int]) -> None:
        current = self._get_tuning()
        current.update(values)
        _check(_lib().mpsb_session_set_tuning(self._h, C.byref(_Tuning(**current))))

    @property
    def tuning(self) -> dict[str, int]:
        """The histogram launch shape (the autotuner's knobs)."""

        return {k: v for k, v in self._get_tuning().items() if k != "planner"}

    @tuning.setter
    def tuning(self, values: dict[str, int]) -> None:
        self._set_tuning(values)

    @property
    def planner(self) -> int:
        """Who plans tree levels for this session's fits: 0 automatic, 1 host, 2 device."""

        return self._get_tuning()["planner"]

    @planner.setter
    def planner(self, value: int) -> None:
        self._set_tuning({"planner": int(value)})

    def __del__(self):
        try:
            if self._h:
                _lib().mpsb_session_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def inject_fault(self, stage: int, countdown: int = 1) -> None:
        _check(_lib().mpsb_debug_inject_fault(self._h, C.c_int32(stage), C.c_int32(countdown)))


FAULT_STAGES = {"off": 0, "upload": 1, "gradients": 2, "histogram": 3, "scan": 4, "partition": 5,
                "prediction": 6, "allocation": 7}


class _Model:
    """RegressionModel / MulticlassModel produced by the device trainer."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        info = _ModelInfo()
        _check(_lib().mpsb_model_info_get(self._h, C.byref(info)))
        self.tree_count = info.tree_count
        self.class_count = info.class_count
        self.feature_count = info.features
        self.learning_rate = info.learning_rate
        self.base_scores = np.array(info.base_scores[: info.class_count], dtype=np.float64)
        self.base_score = float(self.base_scores[0])
        self.max_bins = info.max_bins
        # RegressionModel / MulticlassModel properties (src/python/binding_model.cpp:65-106).
        self.objective = "softmax" if info.class_count > 1 else OBJECTIVE_NAMES[info.objective]
        self.objective_alpha = info.objective_alpha
        self.tweedie_variance_power = info.tweedie_variance_power
        self.class_labels = [float(x) for x in info.class_labels[: info.class_count]] if info.class_count > 1 else []

    def schema(self) -> _BinnedDataset:
        """The frozen quantisation schema (zero rows) for binning prediction inputs."""

        out = C.c_void_p()
        _check(_lib().mpsb_model_schema(self._h, C.byref(out)))
        return _BinnedDataset(out)

    def set_class_labels(self, labels) -> None:
        """MulticlassModel::SetClassLabels: one finite, strictly increasing label per class."""

        arr = np.ascontiguousarray(labels, dtype=np.float64)
        _check(_lib().mpsb_model_set_class_labels(self._h, _p(arr, C.c_double), C.c_uint32(arr.size)))
        self.class_labels = arr.tolist()

    def __del__(self):
        try:
            _lib().mpsb_model_free(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def tree(self, index: int) -> np.ndarray:
        n = C.c_uint32()
        rc = _lib().mpsb_model_tree(self._h, C.c_uint32(index), None, C.c_uint32(0), C.byref(n))
        nodes = np.zeros(n.value, dtype=NODE_DTYPE)
        if rc != 0 and n.value == 0:
            _check(rc)
        _check(_lib().mpsb_model_tree(self._h, C.c_uint32(index), nodes.ctypes.data_as(C.c_void_p), C.c_uint32(n.value), C.byref(n)))
        return nodes

    @property
    def trees(self) -> list[np.ndarray]:
        return [self.tree(i) for i in range(self.tree_count)]

    def fit_stats(self) -> dict[str, float]:
        st = _FitStats()
        _check(_lib().mpsb_model_fit_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in _FitStats._fields_}

    def serialize(self) -> bytes:
        """SerializeModel: the reference's MPSBMOD 1.4 bytes."""

        size = C.c_uint64()
        _check(_lib().mpsb_model_serialize(self._h, None, C.c_uint64(0), C.byref(size)))
        out = (C.c_uint8 * size.value)()
        _check(_lib().mpsb_model_serialize(self._h, out, C.c_uint64(size.value), C.byref(size)))
        return bytes(out)

    def save(self, path: str) -> None:
        """SaveModelFile: atomic write (temporary file, fsync, rename)."""

        _check(_lib().mpsb_model_save(self._h, os.fsencode(path)))

    def predict_binned(self, session: _Session, dataset: _BinnedDataset) -> np.ndarray:
        out = np.zeros(dataset.rows * self.class_count, dtype=np.float64)
        _check(_lib().mpsb_model_predict(session.handle, self._h, dataset.handle, _p(out, C.c_double)))
        return out if self.class_count == 1 else out.reshape(dataset.rows, self.class_count)


def _deserialize_model(data: bytes, kind: int = -1) -> "_Model":
    """DeserializeModel (kind 0) / DeserializeMulticlassModel (kind 1) with the reference's
    validation; kind -1 accepts either."""

    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    out = C.c_void_p()
    _check(_lib().mpsb_model_deserialize_kind(buf, C.c_uint64(len(data)), C.c_int32(kind), C.byref(out)))
    return _Model(out)


def _load_model(path: str, kind: int = -1) -> "_Model":
    """LoadModelFile (kind 0) / LoadMulticlassModelFile (kind 1); kind -1 accepts either."""

    out = C.c_void_p()
    _check(_lib().mpsb_model_load_kind(os.fsencode(path), C.c_int32(kind), C.byref(out)))
    return _Model(out)


def _train_binned(session: _Session, dataset: _BinnedDataset, labels, sample_weights,
                  parameters: _TrainingParameters, class_count: int = 0) -> _Model:
    y = np.ascontiguousarray(labels, dtype=np.float64)
    if y.ndim != 1 or y.size != dataset.rows:
        raise TrainingError("Training data, labels, or max_bins contract does not match")
    w = None if sample_weights is None else np.ascontiguousarray(sample_weights, dtype=np.float64)
    if w is not None and w.size != y.size:
        raise TrainingError("labels and sample weights must have the same length")
    cp, keep = parameters.to_c()
    out = C.c_void_p()
    wp = None if w is None else _p(w, C.c_double)
    if class_count:
        _check(_lib().mpsb_train_multiclass(session.handle, dataset.handle, _p(y, C.c_double), wp,
                                            C.c_uint32(class_count), C.byref(cp), C.byref(out)))
    else:
        _check(_lib().mpsb_train_regressor(session.handle, dataset.handle, _p(y, C.c_double), wp, C.byref(cp),
                                           C.byref(out)))
    del keep
    return _Model(out)


class _RegressionModel(_Model):
    """Python face of a trained or loaded model (the reference's ``_native`` RegressionModel /
    MulticlassModel): ``predict`` / ``predict_margins`` bin the input on the device with the
    model's frozen schema and evaluate every tree there."""

    def __init__(self, model: _Model, schema: _BinnedDataset | None = None, session: _Session | None = None,
                 device: int = 0):
        self.__dict__.update(model.__dict__)
        model._h = None  # ownership moves here
        self._schema = schema
        self._session = session
        self._device = device

    def _inputs(self, matrix):
        if self._schema is None:
            self._schema = self.schema()
        if self._session is None:
            self._session = _session(self._device)
        return _transform_dense_device(self._session, matrix, self._schema)

    def predict(self, matrix) -> np.ndarray:
        """RegressionModel::Predict: raw predictions (margins for binary logistic, log-means
        for Poisson / Tweedie)."""

        data = self._inputs(matrix)
        return self.predict_binned(self._session, data)

    def predict_margins(self, matrix) -> np.ndarray:
        """MulticlassModel::PredictMargins: row-major [rows][class_count], flattened."""

        data = self._inputs(matrix)
        return self.predict_binned(self._session, data).ravel()


_thread_sessions = threading.local()


def _session(device: int = 0) -> _Session:
    """This thread's session on `device`: threads that train concurrently get their own
    sessions (streams), so their kernels overlap on the GPU."""

    sessions = getattr(_thread_sessions, "by_device", None)
    if sessions is None:
        sessions = _thread_sessions.by_device = {}
    if device not in sessions:
        sessions[device] = _Session(device)
    return sessions[device]


# Set by mpsboost_amd.autotune.enable(): called as hook(session, dataset, parameters) before each
# fit to apply the launch shape tuned for the dataset's shape class.
_tuning_hook = None


def _apply_tuning(session: _Session, dataset: _BinnedDataset, parameters: _TrainingParameters) -> None:
    if _tuning_hook is not None:
        _tuning_hook(session, dataset, parameters)


def _train_regressor_hip(matrix, labels, sample_weights, parameters: _TrainingParameters, device: int = 0):
    """TrainRegressorMps (src/python/binding_model.cpp:34-45) on the MI355X backend: binning on
    the device, the codes left resident for the fit."""

    y = np.ascontiguousarray(labels, dtype=np.float64)
    w = None if sample_weights is None else np.ascontiguousarray(sample_weights, dtype=np.float64)
    if w is not None and np.all(w == 1.0):
        w = None  # all-ones weights: the unit-Hessian path, identical arithmetic
    s = _session(device)
    ds = _quantize_dense_device(s, matrix, parameters.max_bins)
    _apply_tuning(s, ds, parameters)
    # The model bins prediction inputs with its own frozen schema (no codes), so the training
    # codes, which stay on the device, are never copied to the host.
    return _RegressionModel(_train_binned(s, ds, y, w, parameters), None, s, device)


def _train_multiclass_softmax_hip(matrix, labels, sample_weights, parameters: _TrainingParameters,
                                  class_count: int, device: int = 0):
    """TrainMulticlassSoftmaxCpu's device counterpart (src/python/binding_model.cpp:47-59)."""

    s = _session(device)
    ds = _quantize_dense_device(s, matrix, parameters.max_bins)
    _apply_tuning(s, ds, parameters)
    return _RegressionModel(_train_binned(s, ds, labels, sample_weights, parameters, class_count), None, s, device)


def _load_regression_model(path: str, device: int = 0) -> _RegressionModel:
    """LoadModelFile (src/io/model_file_io.cpp), bound to the device for prediction."""

    return _RegressionModel(_load_model(path, 0), device=device)


def _load_multiclass_model(path: str, device: int = 0) -> _RegressionModel:
    """LoadMulticlassModelFile (src/io/model_file_io.cpp), bound to the device for prediction."""

    return _RegressionModel(_load_model(path, 1), device=device)


# ----------------------------------------------------------------- test hooks -------
class _HipBackend:
    """Reusable device test session with the reference ``_MpsBackend`` hook methods
    (src/python/binding_backend.cpp:19-84)."""

    def __init__(self, device: int = 0):
        self._session = _Session(device)

    @property
    def session(self) -> _Session:
        return self._session

    def gradients(self, labels, predictions):
        y = np.ascontiguousarray(labels, dtype=np.float64)
        p = np.ascontiguousarray(predictions, dtype=np.float64)
        if y.size != p.size:
            raise TrainingError("Label and prediction lengths do not match")
        out = np.zeros((y.size, 2), dtype=np.float64)
        _check(_lib().mpsb_compute_squared_error(self._session.handle, _p(y, C.c_double), _p(p, C.c_double),
                                                 C.c_uint64(y.size), _p(out, C.c_double)))
        return [tuple(r) for r in out.tolist()]

    @staticmethod
    def _pairs(dataset, labels, predictions):
        y = np.ascontiguousarray(labels, dtype=np.float64)
        p = np.ascontiguousarray(predictions, dtype=np.float64)
        if y.size != dataset.rows or p.size != dataset.rows:
            raise TrainingError("Gradient count does not match binned dataset row count")
        return np.ascontiguousarray(np.stack([p - y, np.ones_like(y)], axis=1))

    def histogram_arrays(self, dataset: _BinnedDataset, pairs, rows, precision: str = "fp64"):
        """Reference-view (count, G, H) arrays of one node from FP64 (g, h) pairs."""

        return self.layer_histogram_arrays(dataset, pairs, [rows], precision)[0]

    def layer_histogram_arrays(self, dataset: _BinnedDataset, pairs, node_rows, precision: str = "fp64"):
        gh = np.ascontiguousarray(pairs, dtype=np.float64).reshape(-1, 2)
        if gh.shape[0] != dataset.rows:
            raise TrainingError("Gradient count does not match binned dataset row count")
        lists = [np.ascontiguousarray(r, dtype=np.uint64) for r in node_rows]
        offsets = np.zeros(len(lists) + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum([r.size for r in lists])
        rows = np.concatenate(lists) if lists else np.zeros(0, dtype=np.uint64)
        cells = int(dataset.bin_counts.sum())
        n = len(lists)
        cnt = np.zeros(n * cells, dtype=np.uint64)
        g = np.zeros(n * cells, dtype=np.float64)
        h = np.zeros(n * cells, dtype=np.float64)
        _check(
            _lib().mpsb_build_layer_histograms(
                self._session.handle, dataset.handle, C.c_uint32(n), _p(offsets, C.c_uint64),
                _p(rows, C.c_uint64), _p(gh, C.c_double), C.c_int32(PRECISIONS[precision]),
                _p(cnt, C.c_uint64), _p(g, C.c_double), _p(h, C.c_double),
            )
        )
        return [(cnt[i * cells:(i + 1) * cells], g[i * cells:(i + 1) * cells], h[i * cells:(i + 1) * cells]) for i in range(n)]

    def baseline_arrays(self, dataset: _BinnedDataset, pairs, rows):
        gh = np.ascontiguousarray(pairs, dtype=np.float64).reshape(-1, 2)
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        cells = int(dataset.bin_counts.sum())
        cnt = np.zeros(cells, dtype=np.uint64)
        g = np.zeros(cells, dtype=np.float64)
        h = np.zeros(cells, dtype=np.float64)
        _check(
            _lib().mpsb_build_baseline_histograms(
                self._session.handle, dataset.handle, _p(r, C.c_uint64), C.c_uint64(r.size), _p(gh, C.c_double),
                _p(cnt, C.c_uint64), _p(g, C.c_double), _p(h, C.c_double),
            )
        )
        return cnt, g, h

    @staticmethod
    def _to_lists(dataset, cnt, g, h):
        out, at = [], 0
        for f in range(dataset.features):
            nb = int(dataset.bin_counts[f])
            out.append([
                {"count": int(cnt[i]), "gradient_sum": float(g[i]), "hessian_sum": float(h[i])}
                for i in range(at, at + nb)
            ])
            at += nb
        return out

    def histograms(self, dataset, labels, predictions, rows, precision: str = "fp32"):
        """Two-stage histogram (MpsBackend::BuildHistograms test hook, RunMpsHistogramForTest)."""

        cnt, g, h = self.histogram_arrays(dataset, self._pairs(dataset, labels, predictions), rows, precision)
        t = self.last_timing
        return {
            "histograms": self._to_lists(dataset, cnt, g, h),
            "encode_seconds": t["histogram_encode_seconds"],
            "command_seconds": t["histogram_command_seconds"],
        }

    def baseline_histograms(self, dataset, labels, predictions, rows):
        cnt, g, h = self.baseline_arrays(dataset, self._pairs(dataset, labels, predictions), rows)
        return self._to_lists(dataset, cnt, g, h)

    def split_candidates(self, dataset, labels, predictions, rows, min_samples_leaf=1, min_child_weight=0.0,
                         reg_lambda=1.0, gamma=0.0, reg_alpha=0.0, min_gain_to_split=0.0, precision="fp64",
                         pairs=None):
        """Per-feature best candidates (ScanSplitsForTest) plus the node's best as ``best``."""

        gh = self._pairs(dataset, labels, predictions) if pairs is None else np.ascontiguousarray(pairs, dtype=np.float64)
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        sp = _SplitParams(min_samples_leaf, min_child_weight, reg_lambda, reg_alpha, gamma, min_gain_to_split)
        out = (_SplitCandidate * dataset.features)()
        best = _SplitCandidate()
        _check(
            _lib().mpsb_scan_splits(
                self._session.handle, dataset.handle, _p(r, C.c_uint64), C.c_uint64(r.size), _p(gh, C.c_double),
                C.byref(sp), C.c_int32(PRECISIONS[precision]), out, C.byref(best),
            )
        )

        def as_dict(c):
            d = {name: getattr(c, name) for name, _ in _SplitCandidate._fields_}
            d["valid"] = bool(d["valid"])
            d["default_left"] = bool(d["default_left"])
            return d

        cands = [as_dict(c) for c in out]
        self.best = as_dict(best)
        return cands

    def partition_rows(self, dataset, rows, feature, threshold_bin, default_left=True):
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        left = np.zeros(max(1, r.size), dtype=np.uint64)
        right = np.zeros(max(1, r.size), dtype=np.uint64)
        nl, nr = C.c_uint64(), C.c_uint64()
        _check(
            _lib().mpsb_partition_rows(
                self._session.handle, dataset.handle, _p(r, C.c_uint64), C.c_uint64(r.size), C.c_uint32(feature),
                C.c_uint32(threshold_bin), C.c_int32(1 if default_left else 0), _p(left, C.c_uint64),
                C.byref(nl), _p(right, C.c_uint64), C.byref(nr),
            )
        )
        return left[: nl.value].tolist(), right[: nr.value].tolist()

    @property
    def last_timing(self) -> dict[str, float]:
        t = _Timing()
        _check(_lib().mpsb_session_last_timing(self._session.handle, C.byref(t)))
        return {name: getattr(t, name) for name, _ in _Timing._fields_}


def _vector_add(left, right, device: int = 0):
    a = np.ascontiguousarray(left, dtype=np.float32)
    b = np.ascontiguousarray(right, dtype=np.float32)
    if a.size != b.size:
        raise ValueError("vector lengths must match")
    out = np.zeros_like(a)
    _check(_lib().mpsb_vector_add(_session(device).handle, _p(a, C.c_float), _p(b, C.c_float), C.c_uint64(a.size), _p(out, C.c_float)))
    return out.tolist()
