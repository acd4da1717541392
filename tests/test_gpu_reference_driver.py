"""The drop-in boundary exercised by the reference's own code.

oracle/_ref/libref_hip.so is the reference's training core (src/core/trainer.cpp,
multiclass_trainer.cpp, tree.cpp, tree_histogram.cpp, tree_split.cpp ...) compiled from its
sources with include/mpsboost_hip_backend.hpp's HipBackend in place of CpuReferenceBackend:
its gradients and (layer) histograms come from the device through the C ABI, everything else
is the reference's host code. In the row-order FP64 form the device returns the reference's
own sums, so the result must be bit-identical to the same trainers on the CPU backend
(oracle/_ref/libref.so).
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

REF = O.reference()
REF_HIP = O.reference_hip()
needs_both = pytest.mark.skipif(REF is None or REF_HIP is None,
                                reason="oracle/_ref was not built (reference sources absent at build time)")

FIELDS = ["feature_index", "threshold_bin", "left_child", "right_child", "default_left", "flags", "leaf_value",
          "gain"]


def _data(seed, rows, feats):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    X[:, 1] = rng.integers(0, 7, rows)
    X[rng.random(X.shape) < 0.06] = np.nan
    y = np.sin(np.nan_to_num(X[:, 0])) + np.nan_to_num(X[:, 2]) * np.nan_to_num(X[:, 3]) + 0.1 * rng.standard_normal(rows)
    return X, y


def _identical(a, b):
    assert a.tree_count == b.tree_count
    for i, (ta, tb) in enumerate(zip(a.trees, b.trees)):
        assert ta.size == tb.size, f"tree {i}: node count"
        for k in FIELDS:
            assert np.array_equal(ta[k], tb[k]), f"tree {i}: {k}"


@needs_both
def test_reference_histograms_through_adapter():
    X, y = _data(3, 2500, 7)
    dc, dh = REF.quantize(X, 32), REF_HIP.quantize(X, 32)
    pairs = REF.squared_error(y, np.zeros_like(y))
    rows = np.arange(1, 2500, 2)
    for a, b in zip(REF.histograms(dc, pairs, rows), REF_HIP.histograms(dh, pairs, rows)):
        assert np.array_equal(a, b)


@needs_both
def test_reference_regressor_on_hip_backend():
    X, y = _data(5, 4000, 8)
    dc, dh = REF.quantize(X, 64), REF_HIP.quantize(X, 64)
    kw = dict(n_estimators=5, learning_rate=0.3, max_depth=5, min_samples_leaf=3, reg_alpha=0.1)
    a, b = REF.train_regressor(dc, y, **kw), REF_HIP.train_regressor(dh, y, **kw)
    _identical(a, b)
    assert np.array_equal(a.predict(dc), b.predict(dh))


@needs_both
def test_reference_multiclass_on_hip_backend():
    X, _ = _data(7, 3000, 6)
    y = np.random.default_rng(8).integers(0, 3, X.shape[0]).astype(np.float64)
    dc, dh = REF.quantize(X, 32), REF_HIP.quantize(X, 32)
    kw = dict(n_estimators=2, learning_rate=0.3, max_depth=4, min_samples_leaf=5)
    a, b = REF.train_multiclass(dc, y, 3, **kw), REF_HIP.train_multiclass(dh, y, 3, **kw)
    _identical(a, b)
