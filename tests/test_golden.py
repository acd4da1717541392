"""Golden vectors written by the reference's own compiled sources (tests/golden/make_golden.py):
the oracle, the host binning and the device's row-order FP64 form must reproduce them exactly.
They travel with the repository, so these checks need neither the reference tree nor its
compiled checker."""

import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))
FIELDS = ["feature_index", "threshold_bin", "left_child", "right_child", "default_left", "flags", "leaf_value",
          "gain"]

# Must match tests/golden/make_golden.py TREE_CASES.
TREE_CASES = [
    ("plain", 32, 3, 0.3, dict(max_depth=5, min_samples_leaf=3)),
    ("regularised", 32, 3, 0.3, dict(max_depth=4, reg_lambda=0.5, reg_alpha=0.2, max_delta_step=0.6, gamma=0.05,
                                      min_child_weight=2.0)),
    ("constrained", 16, 3, 0.5, dict(max_depth=4, monotonic_constraints=[1, 0, -1, 0, 0, 0],
                                      interaction_constraints=[[0, 1, 2], [3, 4, 5]], min_samples_leaf=4)),
    ("random_threshold", 32, 2, 0.3, dict(max_depth=3, split_strategy="random_threshold", random_seed=11)),
]


def _trees(prefix):
    nodes, sizes = G[f"{prefix}_nodes"], G[f"{prefix}_sizes"]
    return np.split(nodes, np.cumsum(sizes)[:-1])


def _same_trees(got, want):
    assert len(got) == len(want)
    for i, (a, b) in enumerate(zip(got, want)):
        assert a.size == b.size, f"tree {i}: node count"
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), f"tree {i}: {k}"


def _check_quantized(ds, mb):
    assert np.array_equal(ds.bins, G[f"q{mb}_bins"])
    assert np.array_equal(ds.missing.astype(np.uint8), G[f"q{mb}_missing"])
    assert np.array_equal(ds.bin_counts, G[f"q{mb}_bin_counts"])
    flat = np.concatenate([np.asarray(b, dtype=np.float32) for b in ds.boundaries])
    assert np.array_equal(flat, G[f"q{mb}_boundaries"])


@pytest.mark.parametrize("mb", [16, 32, 300])
def test_oracle_quantize_golden(mb):
    _check_quantized(O.oracle().quantize(G["X"], mb), mb)


@pytest.mark.parametrize("mb", [16, 32, 300])
def test_host_quantize_golden(mb):
    N = pytest.importorskip("mpsboost_amd._native")
    _check_quantized(N._quantize_dense(G["X"], mb, threads=4), mb)


def test_oracle_histogram_golden():
    o = O.oracle()
    c, g, h = o.histograms(o.quantize(G["X"], 32), G["pairs"], G["rows_sel"])
    assert np.array_equal(c, G["hist32_count"]) and np.array_equal(g, G["hist32_g"])
    assert np.array_equal(h, G["hist32_h"])


@pytest.mark.parametrize("case", TREE_CASES, ids=[c[0] for c in TREE_CASES])
def test_oracle_trees_golden(case):
    name, mb, n, lr, kw = case
    o = O.oracle()
    ds = o.quantize(G["X"], mb)
    m = o.train_regressor(ds, G["y"], n_estimators=n, learning_rate=lr, **kw)
    _same_trees(m.trees, _trees(f"tree_{name}"))
    assert np.array_equal(m.predict(ds), G[f"tree_{name}_pred"])


def test_oracle_multiclass_golden():
    o = O.oracle()
    ds = o.quantize(G["X"], 32)
    m = o.train_multiclass(ds, G["labels3"], 3, n_estimators=2, learning_rate=0.3, max_depth=3, min_samples_leaf=5)
    _same_trees(m.trees, _trees("multi"))
    assert np.array_equal(m.predict(ds), G["multi_pred"])


# ----------------------------------------------------------------- device, FP64 form ---
def _params(N, mb, n, lr, kw):
    kw = dict(kw)
    depth = kw.pop("max_depth")
    msl = kw.pop("min_samples_leaf", 1)
    mcw = kw.pop("min_child_weight", 0.0)
    lam = kw.pop("reg_lambda", 1.0)
    return N._TrainingParameters(n, lr, mb, depth, msl, mcw, lam, precision="fp64", **kw)


@pytest.mark.gpu
def test_device_histogram_golden():
    N = pytest.importorskip("mpsboost_amd._native")
    be = N._HipBackend()
    ds = N._quantize_dense(G["X"], 32)
    c, g, h = be.histogram_arrays(ds, G["pairs"], G["rows_sel"], "fp64")
    assert np.array_equal(c, G["hist32_count"]) and np.array_equal(g, G["hist32_g"])
    assert np.array_equal(h, G["hist32_h"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", TREE_CASES, ids=[c[0] for c in TREE_CASES])
def test_device_trees_golden(case):
    N = pytest.importorskip("mpsboost_amd._native")
    name, mb, n, lr, kw = case
    session = N._HipBackend().session
    ds = N._quantize_dense(G["X"], mb)
    m = N._train_binned(session, ds, G["y"], None, _params(N, mb, n, lr, kw))
    _same_trees(m.trees, _trees(f"tree_{name}"))
    assert np.array_equal(m.predict_binned(session, ds), G[f"tree_{name}_pred"])


@pytest.mark.gpu
def test_device_multiclass_golden():
    """Softmax probabilities use the device's exp: tree structure must match exactly; leaves
    and margins to the last few bits."""

    N = pytest.importorskip("mpsboost_amd._native")
    session = N._HipBackend().session
    ds = N._quantize_dense(G["X"], 32)
    m = N._train_binned(session, ds, G["labels3"], None,
                        _params(N, 32, 2, 0.3, dict(max_depth=3, min_samples_leaf=5)), class_count=3)
    want = _trees("multi")
    assert m.tree_count == len(want)
    for a, b in zip(m.trees, want):
        for k in FIELDS[:6]:
            assert np.array_equal(a[k], b[k])
        np.testing.assert_allclose(a["leaf_value"], b["leaf_value"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(m.predict_binned(session, ds), G["multi_pred"], rtol=1e-12, atol=1e-15)
