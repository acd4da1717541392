"""Batched softmax rounds (class_rounds="batched"): the Jacobi form of the multiclass trainer.

Per round, every class's softmax gradients come from the round's starting margins, and the K
class trees grow together as one device-planned tree of K roots. This departs from the
reference's class order (src/core/multiclass_trainer.cpp:132-147) by design, so the check is a
replay of the Jacobi form built from pinned pieces: the oracle's tree builder (oracle.train_tree,
pinned to the compiled reference) grows each class tree from the same gradient pairs, and the
margins of all classes move after the round's K trees.

* FP64_FIXED: the pairs are the FP64 softmax gradients; device trees equal the oracle's up to
  near-ties (gains within 1e-9 relative at the first differing node).
* INT16 / INT8: both channels stochastically rounded with one scale per channel over the round's
  K x rows values (max |g|, max |h| of all classes), the device's counter-based stream keyed by
  (round seed, k * rows + row, channel) replayed here; trees equal up to near-ties.
* Bad parameters raise the documented errors.
"""

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_parity import _match_until_near_tie, _params, _round_channel

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")


@pytest.fixture(scope="module")
def backend():
    return N._HipBackend()


def _data(classes, rows=12000, seed=0):
    rng = np.random.default_rng(91 + classes + seed)
    X = rng.standard_normal((rows, 8)).astype(np.float32)
    y = np.clip(np.floor((X[:, 0] + 0.7 * X[:, 1] + 1.5) * classes / 3.0), 0, classes - 1).astype(np.float64)
    noise = rng.random(rows) < 0.2
    y[noise] = rng.integers(0, classes, int(noise.sum()))
    return X, y


def _jacobi_replay(o, ods, y, w, classes, rounds, lr, seed, qmax, kw):
    """The Jacobi form on the oracle: per round, all classes' (g, h) from the same margins
    (exponentials summed in class order, like the device), then one tree per class."""

    rows = y.size
    wv = np.ones(rows) if w is None else w
    tw = np.cumsum(wv)[-1]
    base = [np.log(max(1e-15, np.cumsum(np.where(y == c, wv, 0.0))[-1] / tw)) for c in range(classes)]
    margins = np.tile(np.array(base), (rows, 1))
    trees = []
    for r in range(rounds):
        top = margins.max(axis=1, keepdims=True)
        e = np.exp(margins - top)
        z = np.zeros(rows)
        for k in range(classes):
            z = z + e[:, k]
        g = np.concatenate([(e[:, k] / z - (y == k)) * wv for k in range(classes)])
        h = np.concatenate([(e[:, k] / z) * (1.0 - e[:, k] / z) * wv for k in range(classes)])
        if qmax:
            s = (seed + r * 0x9E3779B97F4A7C15) % (1 << 64)
            g, h = _round_channel(g, s, 0, qmax), _round_channel(h, s, 1, qmax)
        round_trees = []
        for k in range(classes):
            pairs = np.stack([g[k * rows:(k + 1) * rows], h[k * rows:(k + 1) * rows]], axis=1)
            round_trees.append(o.train_tree(ods, np.ascontiguousarray(pairs), **kw))
        for k, nodes in enumerate(round_trees):
            margins[:, k] += lr * o.tree_predict(ods, nodes, ods.features)
        trees += round_trees
    return trees


@pytest.mark.parametrize("form,qmax,classes,weighted", [
    ("fp64_fixed", 0.0, 3, False),
    ("fp64_fixed", 0.0, 6, True),
    ("int16", 32767.0, 3, True),
    ("int16", 32767.0, 10, False),
    ("int8", 127.0, 4, False),
])
def test_batched_rounds_match_jacobi_replay(backend, form, qmax, classes, weighted):
    X, y = _data(classes)
    w = np.random.default_rng(5).uniform(0.3, 2.0, y.size) if weighted else None
    o = O.oracle()
    ods, ds = o.quantize(X, 64), N._quantize_dense(X, 64)
    kw = dict(max_depth=4, min_samples_leaf=10)
    rounds, lr, seed = 3, 0.3, 29
    model = N._train_binned(backend.session, ds, y, w,
                            _params(rounds, lr, 64, precision=form, rounding_seed=seed, leaf_solve="histogram",
                                    near_ties="keep", class_rounds="batched", **kw),
                            class_count=classes)
    ref = _jacobi_replay(o, ods, y, w, classes, rounds, lr, seed, qmax, kw)
    assert model.tree_count == len(ref) == rounds * classes
    matched = _match_until_near_tie(model.trees, ref, ods, 1e-9, f"batched {form} K={classes}")
    print(f"batched {form} K={classes}: {matched} of {len(ref)} trees identical")
    assert matched >= classes, f"only {matched} trees matched before a near-tie"
    st = model.fit_stats()
    assert st["trees"] == rounds * classes and st["device_planned_trees"] == rounds * classes


def test_batched_rounds_renewed_leaves_and_predictions(backend):
    """With leaf renewal the leaves are solved from the round's FP64 pairs; the model's margins
    for the training rows must equal the margins the fit itself carried (prediction walks the
    trees in the reference's round-major class order)."""

    classes = 5
    X, y = _data(classes, rows=20000, seed=3)
    ds = N._quantize_dense(X, 64)
    p = _params(4, 0.2, 64, precision="int16", rounding_seed=3, leaf_solve="renewed", near_ties="keep",
                class_rounds="batched", max_depth=5, min_samples_leaf=5)
    model = N._train_binned(backend.session, ds, y, None, p, class_count=classes)
    assert model.tree_count == 4 * classes
    margins = model.predict_binned(backend.session, ds)
    acc = float(np.mean(np.argmax(margins, axis=1) == y))
    seq = N._train_binned(backend.session, ds, y, None,
                          _params(4, 0.2, 64, precision="int16", rounding_seed=3, leaf_solve="renewed",
                                  near_ties="keep", max_depth=5, min_samples_leaf=5),
                          class_count=classes)
    acc_seq = float(np.mean(np.argmax(seq.predict_binned(backend.session, ds), axis=1) == y))
    print(f"batched accuracy {acc:.5f}, class order {acc_seq:.5f}")
    assert abs(acc - acc_seq) < 0.02
    st = model.fit_stats()
    assert st["renewed_trees"] == 4


@pytest.mark.parametrize("change,message", [
    (dict(precision="fp32"), "fp64_fixed, int16 or int8"),
    (dict(near_ties="resolve"), "keep near ties"),
    (dict(max_leaves=4), "leaf cap"),
    (dict(max_depth=12), "2048"),
])
def test_batched_rounds_reject_unsupported(backend, change, message):
    X, y = _data(3, rows=2000)
    ds = N._quantize_dense(X, 32)
    kw = dict(precision="int16", near_ties="keep", class_rounds="batched", max_depth=4, min_samples_leaf=5)
    kw.update(change)
    with pytest.raises(ValueError, match=message):
        N._train_binned(backend.session, ds, y, None, _params(2, 0.3, 32, **kw), class_count=3)


def test_classifier_batched_softmax_strategy(backend):
    """GradientBoostingClassifier(multi_strategy="batched_softmax"): the device extension reaches
    the batched rounds through the estimator and predicts like the class-order fit."""

    from mpsboost_amd import estimators as E

    X, y = _data(4, rows=8000, seed=5)
    kw = dict(n_estimators=6, max_depth=4, learning_rate=0.3, precision="int16", min_samples_leaf=10)
    a = E.GradientBoostingClassifier(multi_strategy="batched_softmax", **kw).fit(X, y)
    b = E.GradientBoostingClassifier(multi_strategy="softmax", **kw).fit(X, y)
    assert a.training_summary_["class_rounds"] == "batched"
    assert b.training_summary_["class_rounds"] == "class_order"
    assert a.n_estimators_ == b.n_estimators_ == 6 * 4
    acc_a = float(np.mean(a.predict(X) == y))
    acc_b = float(np.mean(b.predict(X) == y))
    assert abs(acc_a - acc_b) < 0.03, (acc_a, acc_b)


@pytest.mark.parametrize("classes,depth,nan_rate,weighted,msl", [
    (2, 1, 0.0, False, 1),
    (3, 2, 0.1, True, 5),
    (5, 6, 0.0, True, 30),
])
def test_batched_rounds_edges(backend, classes, depth, nan_rate, weighted, msl):
    """Two classes, stumps, missing values (their own histogram cell and default direction) and a
    wide tree with a large leaf minimum, FP64_FIXED against the Jacobi replay. (The wide case has
    no missing values: with them, the oracle's row-order FP64 sums meet small softmax Hessians
    and raise the reference's own "missing statistics exceed parent statistics", as
    test_gpu_parity.py's multiclass case shows.)"""

    X, y = _data(classes, rows=9000, seed=11)
    if nan_rate:
        X[np.random.default_rng(12).random(X.shape) < nan_rate] = np.nan
    w = np.random.default_rng(13).uniform(0.5, 1.5, y.size) if weighted else None
    o = O.oracle()
    ods, ds = o.quantize(X, 32), N._quantize_dense(X, 32)
    kw = dict(max_depth=depth, min_samples_leaf=msl)
    rounds, lr = 3, 0.4
    model = N._train_binned(backend.session, ds, y, w,
                            _params(rounds, lr, 32, precision="fp64_fixed", leaf_solve="histogram",
                                    class_rounds="batched", **kw),
                            class_count=classes)
    ref = _jacobi_replay(o, ods, y, w, classes, rounds, lr, 0, 0.0, kw)
    matched = _match_until_near_tie(model.trees, ref, ods, 1e-9, f"batched edges K={classes} depth {depth}")
    assert matched >= classes, f"only {matched} trees matched before a near-tie"
