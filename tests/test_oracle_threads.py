"""The oracle's feature-parallel mode (orc_set_threads) against its one-thread form: binning,
trees, gains, leaves and predictions bit-identical, and the same first error, for several
thread counts. Each feature's loop is the same sequence of operations on any thread and the
results are combined in feature order, so nothing may differ."""

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture()
def checker():
    o = O.oracle()
    yield o
    o.set_threads(1)


def _fit(o, threads, X, y, **kw):
    o.set_threads(threads)
    ds = o.quantize(X, 64)
    m = o.train_regressor(ds, y, n_estimators=3, learning_rate=0.3, **kw)
    return ds.boundaries, ds.bins.copy(), [t.tobytes() for t in m.trees], m.predict(ds)


@pytest.mark.parametrize("kw", [
    dict(max_depth=5, min_samples_leaf=5, min_child_weight=1.0),
    dict(max_depth=4, min_samples_leaf=1, min_child_weight=0.0, monotonic_constraints=[1, 0, -1, 0, 0, 0, 0, 0, 0]),
    dict(max_depth=4, min_samples_leaf=3, min_child_weight=0.5, interaction_constraints=[[0, 1], [2, 3, 4]]),
])
def test_threads_bit_identical(checker, kw):
    rng = np.random.default_rng(77)
    X = rng.standard_normal((5000, 9)).astype(np.float32)
    X[:, 3] = rng.integers(0, 4, 5000)
    X[rng.random(X.shape) < 0.07] = np.nan
    y = np.sin(np.nan_to_num(X[:, 0])) + np.nan_to_num(X[:, 2]) ** 2 + 0.1 * rng.standard_normal(5000)
    one = _fit(checker, 1, X, y, **kw)
    for t in (2, 5, 16):
        other = _fit(checker, t, X, y, **kw)
        assert other[0] == one[0]
        assert np.array_equal(other[1], one[1])
        assert other[2] == one[2]
        assert np.array_equal(other[3], one[3])


def test_threads_report_the_first_bad_value(checker):
    X = np.ones((50, 6), dtype=np.float32)
    X[10, 4] = np.inf
    X[3, 2] = np.inf
    msgs = []
    for t in (1, 4):
        checker.set_threads(t)
        with pytest.raises(O.DataError) as e:
            checker.quantize(X, 16)
        msgs.append(str(e.value))
    assert msgs[0] == msgs[1]


@pytest.mark.timeout(120)
def test_pool_survives_resizes_between_runs(checker):
    """Workers created by a resize wait for the next parallel loop, never for one already
    finished: many resizes, each followed by parallel loops, all complete with equal results."""

    X = np.random.default_rng(3).standard_normal((400, 12)).astype(np.float32)
    expect = checker.quantize(X, 32).bins.copy()
    for k in range(60):
        checker.set_threads(1 + k % 7)
        for _ in range(3):
            assert np.array_equal(checker.quantize(X, 32).bins, expect)
