"""The launch-shape autotuner: session tuning get/set, the search, the per-shape-class cache and
the hook that applies it. Every candidate shape must train the same model in the exact forms
(integer sums do not depend on how rows and cells are split into blocks)."""

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")
from mpsboost_amd import autotune as A  # noqa: E402


def _data(rows=30_000, feats=20, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    y = np.sin(X[:, 0]) + X[:, 1] * X[:, 2] + 0.1 * rng.standard_normal(rows)
    return X, y


def test_session_tuning_round_trip_and_validation():
    s = N._Session(0)
    assert s.tuning == A.DEFAULT and s.planner == 0
    s.tuning = {"block": 256, "words_per_thread": 2}
    assert s.tuning["block"] == 256 and s.tuning["words_per_thread"] == 2
    for bad in ({"block": 384}, {"words_per_thread": 3}, {"lds_budget": 1024}, {"target_blocks": 1},
                {"min_slab_rows": 1}, {"lds_budget": 10 ** 9}, {"planner": 3}):
        with pytest.raises(ValueError, match="tuning"):
            s.tuning = bad
    s.planner = 1
    assert s.planner == 1 and s.tuning["block"] == 256


@pytest.mark.parametrize("precision", ["int16", "fp64_fixed"])
def test_every_candidate_trains_the_same_model(precision):
    X, y = _data()
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, 64)
    p = N._TrainingParameters(3, 0.2, 64, 6, 5, 0.0, 1.0, precision=precision, rounding_seed=3)
    ref = N._train_binned(s, ds, y, None, p).serialize()
    for cand in A.CANDIDATES[::5]:
        s.tuning = cand
        assert N._train_binned(s, ds, y, None, p).serialize() == ref, cand


def test_tune_cache_and_hook(tmp_path):
    X, y = _data(seed=1)
    s = N._Session(0)
    ds = N._quantize_dense_device(s, X, 64)
    p = N._TrainingParameters(2, 0.2, 64, 6, 5, 0.0, 1.0, precision="int16")
    path = str(tmp_path / "tuning.json")
    best = A.tune(s, ds, y, p, candidates=A.CANDIDATES[:6], trees=2, repeats=1, path=path)
    assert len(best["_seconds"]) == 6
    key = A.shape_class(ds, "int16")
    assert A.lookup(key) == {k: v for k, v in best.items() if not k.startswith("_")}
    assert key in json.load(open(path))
    assert s.tuning == A.DEFAULT  # the search leaves the session as it found it
    try:
        A.enable(path)
        from mpsboost_amd import estimators as E

        est = E.GradientBoostingRegressor(n_estimators=2, max_bins=64, precision="int16").fit(X, y)
        assert est.n_estimators_ == 2
        assert N._session().tuning == A.lookup(key)
    finally:
        A.disable()
