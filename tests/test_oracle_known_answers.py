"""Pin the CPU oracle against the reference's own hand-computed known answers.

Every expected value below is a constant from the reference's unit tests (cited per test);
none is recomputed from a formula here. Each case runs against the repository's
restatement (``liboracle.so``) and, where it was built, against the reference sources
compiled unchanged (``oracle/_ref/libref.so``), so a case that passes for the reference but
fails for the restatement points at the restatement.
"""

import math

import numpy as np
import pytest

from oracle import oracle as O

CHECKERS = [pytest.param("oracle", id="restatement")]
if O.reference() is not None:
    CHECKERS.append(pytest.param("reference", id="reference"))


@pytest.fixture(params=CHECKERS)
def chk(request):
    return O.oracle() if request.param == "oracle" else O.reference()


def _tree(chk, matrix, labels, *, max_bins=16, **kw):
    ds = chk.quantize(np.asarray(matrix, dtype=np.float32), max_bins)
    pairs = chk.squared_error(labels, [0.0] * len(labels))
    kw.setdefault("reg_lambda", 0.0)
    return ds, chk.train_tree(ds, pairs, **kw)


def test_histogram_hand_values(chk):
    """tests/unit/test_cpu_tree.py:25-48"""

    m = np.array([[0.0, 10.0], [1.0, 10.0], [2.0, 20.0], [3.0, 20.0]], dtype=np.float32)
    ds = chk.quantize(m, 16)
    pairs = chk.squared_error([0.0, 1.0, 2.0, 3.0], [0.0] * 4)
    cnt, g, h = chk.histograms(ds, pairs, [1, 2, 3])
    assert cnt.tolist() == [0, 1, 1, 1, 1, 2]
    assert g.tolist() == [0.0, -1.0, -2.0, -3.0, -1.0, -5.0]
    assert h.tolist() == [0.0, 1.0, 1.0, 1.0, 1.0, 2.0]


def test_root_split_gain_and_leaves(chk):
    """tests/unit/test_cpu_tree.py:51-71: split at bin 1, gain 2.0, leaves 0 and 2."""

    ds, nodes = _tree(chk, [[0.0], [1.0], [2.0], [3.0]], [0.0, 0.0, 2.0, 2.0], max_depth=1)
    assert len(nodes) == 3
    root = nodes[0]
    assert (root["flags"], root["feature_index"], root["threshold_bin"]) == (0, 0, 1)
    assert (root["left_child"], root["right_child"], root["leaf_value"], root["gain"]) == (1, 2, 0.0, 2.0)
    assert root["default_left"] == 1
    assert nodes[1]["flags"] == 1 and nodes[1]["leaf_value"] == 0.0
    assert nodes[2]["flags"] == 1 and nodes[2]["leaf_value"] == 2.0
    assert chk.tree_predict(ds, nodes).tolist() == [0.0, 0.0, 2.0, 2.0]


def test_depth_two_tree_every_node(chk):
    """tests/unit/test_cpu_tree.py:74-93"""

    labels = [0.0, 0.0, 1.0, 1.0, 2.0, 2.0, 3.0, 3.0]
    ds, nodes = _tree(chk, [[float(v)] for v in range(8)], labels, max_depth=2)
    assert len(nodes) == 7
    assert [(int(n["feature_index"]), int(n["threshold_bin"])) for n in nodes[:3]] == [(0, 3), (0, 1), (0, 5)]
    assert [(int(n["left_child"]), int(n["right_child"])) for n in nodes[:3]] == [(1, 2), (3, 4), (5, 6)]
    assert nodes["leaf_value"][3:].tolist() == [0.0, 1.0, 2.0, 3.0]
    assert chk.tree_predict(ds, nodes).tolist() == labels


def test_tie_break_smaller_feature_then_threshold(chk):
    """tests/unit/test_cpu_tree.py:96-104"""

    _, nodes = _tree(chk, [[0.0, 0.0], [1.0, 1.0], [2.0, 2.0], [3.0, 3.0]], [0.0, 1.0, 1.0, 0.0], max_depth=1)
    assert (nodes[0]["feature_index"], nodes[0]["threshold_bin"]) == (0, 0)


def test_random_threshold_is_seeded(chk):
    """tests/unit/test_cpu_tree.py:106-143"""

    ds = chk.quantize(np.asarray([[float(v)] for v in range(8)], dtype=np.float32), 8)
    pairs = chk.squared_error([0.0, 0.0, 1.0, 1.0, 2.0, 2.0, 3.0, 3.0], [0.0] * 8)
    kw = dict(max_depth=1, reg_lambda=0.0, split_strategy="random_threshold")
    first = chk.train_tree(ds, pairs, random_seed=101, **kw)
    second = chk.train_tree(ds, pairs, random_seed=101, **kw)
    other = chk.train_tree(ds, pairs, random_seed=202, **kw)
    assert first.tobytes() == second.tobytes()
    assert first[0]["threshold_bin"] != other[0]["threshold_bin"]


def test_depth_zero_gamma_and_constant_feature(chk):
    """tests/unit/test_cpu_tree.py:146-161"""

    ds, nodes = _tree(chk, [[1.0], [1.0]], [2.0, 4.0], max_depth=0)
    assert len(nodes) == 1 and nodes[0]["leaf_value"] == 3.0
    assert chk.tree_predict(ds, nodes).tolist() == [3.0, 3.0]
    _, blocked = _tree(chk, [[0.0], [1.0], [2.0], [3.0]], [0.0, 0.0, 2.0, 2.0], max_depth=2, gamma=3.0)
    assert len(blocked) == 1 and blocked[0]["leaf_value"] == 1.0


def test_min_child_constraints_deterministic(chk):
    """tests/unit/test_cpu_tree.py:164-179"""

    ds = chk.quantize(np.asarray([[0.0], [1.0], [2.0], [3.0]], dtype=np.float32), 4)
    pairs = chk.squared_error([0.0, 0.0, 0.0, 10.0], [0.0] * 4)
    kw = dict(max_depth=1, min_samples_leaf=2, min_child_weight=2.0, reg_lambda=0.0)
    a, b = chk.train_tree(ds, pairs, **kw), chk.train_tree(ds, pairs, **kw)
    assert a.tobytes() == b.tobytes()
    assert a[0]["threshold_bin"] == 1


def test_invalid_contracts_fail(chk):
    """tests/unit/test_cpu_tree.py:182-227"""

    ds = chk.quantize(np.ones((2, 1), dtype=np.float32), 4)
    pairs = chk.squared_error([1.0, 2.0], [0.0, 0.0])
    with pytest.raises(O.TrainingError, match="min_samples_leaf"):
        chk.train_tree(ds, pairs, max_depth=1, min_samples_leaf=0)
    with pytest.raises(O.TrainingError, match="reg_lambda"):
        chk.train_tree(ds, pairs, max_depth=1, reg_lambda=-1.0)
    with pytest.raises(O.TrainingError, match="min_gain_to_split"):
        chk.train_tree(ds, pairs, max_depth=1, min_gain_to_split=-1.0)
    tree = chk.train_tree(ds, pairs, max_depth=0, reg_lambda=0.0)
    two = chk.quantize(np.ones((2, 2), dtype=np.float32), 4)
    with pytest.raises(O.TrainingError, match="feature count"):
        chk.tree_predict(two, tree, tree_features=1)


def test_missing_default_direction(chk):
    """tests/unit/test_cpu_tree_missing.py:23-47: NaN row joins the right (label 10) side."""

    ds, nodes = _tree(chk, [[0.0], [1.0], [2.0], [np.nan]], [0.0, 0.0, 10.0, 10.0], max_bins=4,
                      max_depth=1, min_samples_leaf=1, min_child_weight=0.0)
    assert nodes[0]["flags"] == 0 and nodes[0]["default_left"] == 0
    assert ds.missing.tolist() == [[False, False, False, True]]
    assert chk.tree_predict(ds, nodes).tolist() == [0.0, 0.0, 10.0, 10.0]


def test_objective_hand_values(chk):
    """tests/unit/test_objective.py:15-29 and tests/unit/test_regularization_controls.py:9-17"""

    assert chk.squared_error([1.0, -2.0, 4.0], [1.5, -3.0, 2.0]).tolist() == [[0.5, 1.0], [-1.0, 1.0], [-2.0, 1.0]]
    assert chk.node_score(-4.0, 2.0, 1.0) == pytest.approx(16.0 / 3.0)
    assert chk.leaf_weight(-4.0, 2.0, 1.0) == pytest.approx(4.0 / 3.0)
    assert chk.split_gain(0.0, 2.0, -4.0, 2.0, 0.0, 0.25) == 1.75
    assert chk.node_score(-4.0, 2.0, 1.0, reg_alpha=1.0) == pytest.approx(3.0)
    assert chk.leaf_weight(-4.0, 2.0, 1.0, reg_alpha=1.0) == pytest.approx(1.0)
    assert chk.leaf_weight(-4.0, 2.0, 1.0, reg_alpha=1.0, max_delta_step=0.5) == 0.5
    assert chk.leaf_weight(0.5, 2.0, 1.0, reg_alpha=1.0) == 0.0
    assert chk.split_gain(0.0, 2.0, -4.0, 2.0, 0.0, 0.25, reg_alpha=1.0) == 0.875


def test_other_objective_hand_values(chk):
    """tests/unit/test_objective.py:32-84"""

    lg = chk.gradients("binary_logistic", [0.0, 1.0, 1.0], [0.0, 2.0, -2.0])
    assert lg[0] == pytest.approx((0.5, 0.25))
    assert lg[1] == pytest.approx((-0.11920292202211755, 0.10499358540350662))
    assert chk.gradients("quantile", [0.0, 2.0], [1.0, 1.0], alpha=0.25).tolist() == [[0.75, 1.0], [-0.25, 1.0]]
    po = chk.gradients("poisson", [0.0, 2.0], [0.0, math.log(2.0)])
    assert po[1] == pytest.approx((0.0, 2.0))
    tw = chk.gradients("tweedie", [0.0, 4.0], [0.0, math.log(4.0)], alpha=1.5)
    assert tw[0] == pytest.approx((1.0, 0.5)) and tw[1] == pytest.approx((0.0, 2.0))
    sm = chk.gradients("softmax", [0.0, 2.0], [2.0, 1.0, 0.0, 0.0, 1.0, 2.0], class_count=3, target_class=2)
    assert sm[0][0] > 0 and sm[1][0] < 0 and sm[0][1] > 0 and sm[1][1] > 0
    assert chk.gradients("binary_logistic", [0.0, 1.0], [-1000.0, 1000.0]).tolist() == [[0.0, 0.0], [0.0, 0.0]]


@pytest.mark.parametrize(
    "call, message",
    [
        (lambda c: c.squared_error([], []), "Labels must not be empty"),
        (lambda c: c.squared_error([math.nan], [0.0]), "must be finite"),
        (lambda c: c.node_score(1.0, -1.0, 1.0), "Hessian"),
        (lambda c: c.leaf_weight(1.0, 0.0, 0.0), "finite and positive"),
        (lambda c: c.split_gain(0.0, 0.0, 1.0, 1.0, 1.0, 0.0), "strictly positive"),
        (lambda c: c.split_gain(1.0, 1.0, 1.0, 1.0, 0.0, -1.0), "gamma"),
    ],
)
def test_invalid_objective_inputs(chk, call, message):
    """tests/unit/test_objective.py:87-111"""

    with pytest.raises(O.TrainingError, match=message):
        call(chk)


def test_quantization_hand_cases(chk):
    """tests/unit/test_quantization.py:16-67"""

    m = np.array([[1.0, 10.0], [2.0, 10.0], [3.0, 20.0], [4.0, 20.0]], dtype=np.float32)
    ds = chk.quantize(m, 2)
    assert ds.boundaries == [[2.0], [10.0]]
    assert ds.bins.tolist() == [[0, 0, 1, 1], [0, 0, 1, 1]]
    assert ds.bin_width == 8
    ds = chk.quantize(np.array([[5.0, 1.0], [5.0, 1.0], [5.0, 2.0], [5.0, 2.0]]), 16)
    assert ds.boundaries[0] == [] and ds.bins[0].tolist() == [0, 0, 0, 0] and ds.boundaries[1] == [1.0]
    m = np.array([[1.0, np.nan], [2.0, 10.0], [np.nan, 20.0], [4.0, np.nan]], dtype=np.float32)
    ds = chk.quantize(m, 3)
    assert ds.boundaries == [[1.0, 2.0], [10.0]]
    assert ds.missing.tolist() == [[False, False, True, False], [True, False, False, True]]
    assert ds.bins[0][2] == 0 and ds.bins[1][0] == 0
    skew = np.concatenate([np.zeros((999, 1), np.float32), np.ones((1, 1), np.float32)])
    ds = chk.quantize(skew, 256)
    assert ds.boundaries == [[0.0]]
    assert (ds.bins[0] == 0).sum() == 999 and (ds.bins[0] == 1).sum() == 1
    for mb, width in [(256, 8), (257, 16), (65536, 16)]:
        assert chk.quantize(np.arange(8, dtype=np.float32).reshape(4, 2), mb).bin_width == width
    strided = np.arange(24, dtype=np.float32).reshape(4, 6)[:, ::2]
    assert chk.quantize(strided, 4).bins.tolist() == chk.quantize(np.ascontiguousarray(strided), 4).bins.tolist()


@pytest.mark.parametrize(
    "matrix, message",
    [
        (np.array([[np.inf]], dtype=np.float64), "Inf"),
        (np.array([[np.finfo(np.float64).max]], dtype=np.float64), "finite float32 range"),
    ],
)
def test_quantization_rejects_bad_values(chk, matrix, message):
    """tests/unit/test_quantization.py:86-98"""

    with pytest.raises(O.DataError, match=message):
        chk.quantize(matrix, 4)


def test_boosting_reduces_error_and_is_deterministic(chk):
    """tests/unit/test_trainer.py:14-35"""

    X = np.arange(32, dtype=np.float32).reshape(-1, 1)
    y = np.where(X[:, 0] < 16, -2.0, 3.0)
    ds = chk.quantize(X, 16)
    kw = dict(n_estimators=4, learning_rate=0.5, max_depth=2, min_samples_leaf=1, min_child_weight=1.0, reg_lambda=0.0)
    a, b = chk.train_regressor(ds, y, **kw), chk.train_regressor(ds, y, **kw)
    assert np.mean((y - a.predict(ds)) ** 2) < np.mean((y - y.mean()) ** 2) * 0.02
    assert np.array_equal(a.predict(ds), b.predict(ds))
    assert a.tree_count == 4


def test_frozen_boundaries_for_new_data(chk):
    """tests/unit/test_trainer.py:38-54"""

    X = np.array([[0.0], [1.0], [2.0], [3.0]], dtype=np.float32)
    ds = chk.quantize(X, 4)
    model = chk.train_regressor(ds, [0.0, 0.0, 10.0, 10.0], n_estimators=1, learning_rate=1.0, max_depth=1,
                                min_samples_leaf=1, min_child_weight=1.0, reg_lambda=0.0)
    new = chk.transform(np.array([[-100.0], [100.0]], dtype=np.float32), ds)
    p = model.predict(new)
    assert p[0] < p[1]
