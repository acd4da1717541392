"""Pin the CPU oracle against the compiled reference and against brute force.

* Restatement vs reference (``oracle/_ref/libref.so``, the reference's own sources built
  unchanged): bit-for-bit equality of boundaries, bins, histograms, every node of every
  tree and every prediction on seeded inputs with missing values, ties, skewed and
  constant columns, uint8 and uint16 storage, constraints and every objective.
* Brute force that needs neither: the row-order FP64 histogram equals ``numpy.bincount``
  with weights (which also sums in index order), and on tiny inputs the chosen root split
  equals an exhaustive search over (feature, threshold, missing direction) written from
  the gain's closed form.
"""

import itertools

import numpy as np
import pytest

from oracle import oracle as O

REF = O.reference()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref/libref.so not built (reference sources absent)")


def _data(seed, rows, feats, nan_rate=0.1, discrete=False):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    if discrete:
        X[:, : feats // 2] = rng.integers(0, 5, (rows, feats // 2))
    X[:, -1] = 3.0  # constant column: one bin
    X[rng.random(X.shape) < nan_rate] = np.nan
    y = np.nan_to_num(np.sin(X[:, 0])) + np.nan_to_num(X[:, 1]) * np.nan_to_num(X[:, 2]) + 0.1 * rng.standard_normal(rows)
    return X, y


def _same_models(a, b):
    assert a.tree_count == b.tree_count
    assert np.array_equal(a.base_scores, b.base_scores)
    for ta, tb in zip(a.trees, b.trees):
        assert ta.tobytes() == tb.tobytes()


@needs_ref
@pytest.mark.parametrize("max_bins", [2, 16, 255, 256, 257])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_quantize_bit_identical(max_bins, dtype):
    X, _ = _data(11 + max_bins, 777, 5, discrete=True)
    a, b = O.oracle().quantize(X.astype(dtype), max_bins), REF.quantize(X.astype(dtype), max_bins)
    assert a.bin_width == b.bin_width
    assert a.boundaries == b.boundaries
    assert np.array_equal(a.bins, b.bins) and np.array_equal(a.missing, b.missing)
    assert np.array_equal(a.missing_counts, b.missing_counts)


@needs_ref
def test_histograms_bit_identical_on_subset():
    X, y = _data(5, 1025, 4)
    o = O.oracle()
    for mb in (16, 257):
        da, db = o.quantize(X, mb), REF.quantize(X, mb)
        pairs = o.squared_error(y, np.zeros_like(y))
        rows = np.arange(3, 1025, 2)
        for x, z in zip(o.histograms(da, pairs, rows), REF.histograms(db, pairs, rows)):
            assert np.array_equal(x, z)


@needs_ref
@pytest.mark.parametrize(
    "kw",
    [
        dict(max_depth=5, min_samples_leaf=3),
        dict(max_depth=4, reg_lambda=0.0, reg_alpha=0.3, max_delta_step=0.4, gamma=0.01),
        dict(max_depth=6, max_leaves=11, min_child_weight=4.0),
        dict(max_depth=4, monotonic_constraints=[1, -1, 0, 0, 0, 0]),
        dict(max_depth=4, interaction_constraints=[[0, 1], [2, 3, 4]]),
        dict(max_depth=3, split_strategy="random_threshold", random_seed=7),
        dict(max_depth=5, min_gain_to_split=0.5),
    ],
)
def test_regressor_bit_identical(kw):
    X, y = _data(21, 1500, 6, discrete=True)
    o = O.oracle()
    da, db = o.quantize(X, 32), REF.quantize(X, 32)
    a = o.train_regressor(da, y, n_estimators=6, learning_rate=0.3, **kw)
    b = REF.train_regressor(db, y, n_estimators=6, learning_rate=0.3, **kw)
    _same_models(a, b)
    assert np.array_equal(a.predict(da), b.predict(db))


@needs_ref
@pytest.mark.parametrize("objective", ["binary_logistic", "quantile", "poisson", "tweedie"])
def test_other_objectives_bit_identical(objective):
    X, y = _data(31, 800, 4)
    if objective == "binary_logistic":
        y = (y > 0).astype(np.float64)
    elif objective in ("poisson", "tweedie"):
        y = np.abs(y)
    o = O.oracle()
    da, db = o.quantize(X, 16), REF.quantize(X, 16)
    w = np.linspace(0.5, 1.5, y.size)
    kw = dict(n_estimators=4, learning_rate=0.2, objective=objective, objective_alpha=0.7, max_depth=3)
    _same_models(o.train_regressor(da, y, w, **kw), REF.train_regressor(db, y, w, **kw))


@needs_ref
def test_multiclass_bit_identical():
    X, _ = _data(41, 900, 5)
    rng = np.random.default_rng(3)
    y = rng.integers(0, 4, X.shape[0]).astype(np.float64)
    o = O.oracle()
    da, db = o.quantize(X, 16), REF.quantize(X, 16)
    kw = dict(n_estimators=3, learning_rate=0.3, max_depth=3, min_samples_leaf=5)
    a, b = o.train_multiclass(da, y, 4, **kw), REF.train_multiclass(db, y, 4, **kw)
    _same_models(a, b)
    assert np.array_equal(a.predict(da), b.predict(db))


@needs_ref
def test_transform_bit_identical():
    X, _ = _data(51, 600, 4)
    Z, _ = _data(52, 300, 4)
    o = O.oracle()
    ta = o.transform(Z, o.quantize(X, 64))
    tb = REF.transform(Z, REF.quantize(X, 64))
    assert np.array_equal(ta.bins, tb.bins) and np.array_equal(ta.missing, tb.missing)


def test_histogram_equals_row_order_bincount():
    """FP64 bincount with weights accumulates in index order, as the row-outer CPU loop does."""

    X, y = _data(61, 2000, 5)
    o = O.oracle()
    ds = o.quantize(X, 64)
    pairs = o.squared_error(y, np.full_like(y, 0.25))
    rows = np.sort(np.random.default_rng(0).choice(2000, 1300, replace=False))
    cnt, g, h = o.histograms(ds, pairs, rows)
    at = 0
    for f in range(ds.features):
        nb = int(ds.bin_counts[f])
        b = ds.bins[f][rows]
        assert np.array_equal(cnt[at : at + nb], np.bincount(b, minlength=nb))
        assert np.array_equal(g[at : at + nb], np.bincount(b, weights=pairs[rows, 0], minlength=nb))
        assert np.array_equal(h[at : at + nb], np.bincount(b, weights=pairs[rows, 1], minlength=nb))
        at += nb


def _exhaustive_root(ds, pairs, lam, min_leaf):
    """Best root split by brute force over (feature, threshold, missing side), from
    gain = 1/2 (GL^2/(HL+l) + GR^2/(HR+l) - G^2/(H+l)); ties resolve to the first in
    (feature, threshold, missing-left-first) order."""

    best = None
    for f in range(ds.features):
        miss = ds.missing[f]
        for t in range(int(ds.bin_counts[f]) - 1):
            for dl in (True, False):
                left = np.where(miss, dl, ds.bins[f] <= t)
                if left.sum() < min_leaf or (~left).sum() < min_leaf:
                    continue
                gl, hl = pairs[left, 0].sum(), pairs[left, 1].sum()
                gr, hr = pairs[~left, 0].sum(), pairs[~left, 1].sum()
                gain = 0.5 * (gl * gl / (hl + lam) + gr * gr / (hr + lam) - (gl + gr) ** 2 / (hl + hr + lam))
                if gain > 1e-12 and (best is None or gain > best[0] + 1e-9):
                    best = (gain, f, t, dl)
    return best


@pytest.mark.parametrize("seed", range(6))
def test_root_split_matches_exhaustive_search(seed):
    rng = np.random.default_rng(100 + seed)
    X = rng.integers(0, 6, (40, 3)).astype(np.float32)
    X[rng.random(X.shape) < 0.15] = np.nan
    y = rng.standard_normal(40)
    o = O.oracle()
    ds = o.quantize(X, 8)
    pairs = o.squared_error(y, np.zeros(40))
    nodes = o.train_tree(ds, pairs, max_depth=1, reg_lambda=1.0, min_samples_leaf=2)
    best = _exhaustive_root(ds, pairs, 1.0, 2)
    if best is None:
        assert len(nodes) == 1
        return
    root = nodes[0]
    assert (int(root["feature_index"]), int(root["threshold_bin"]), bool(root["default_left"])) == best[1:]
    assert root["gain"] == pytest.approx(best[0], rel=1e-12)
