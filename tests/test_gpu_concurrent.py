"""Concurrent device sessions: regression fits on two threads at once, each thread on its own
session (one HIP stream each), give model bytes identical to the same fits run one after
another. The reference fans independent fits out over threads the same way
(src/mpsboost/estimators/forest.py:136-144, n_jobs).

One thread's session plans every level on the host, so each of its levels ends in a host wait
for the device-published split candidates (the bounded wait of mpsboost_amd/csrc/host_wait.hpp)
while the other thread's session plans on the device: the two planners run side by side, and
the fit statistics show which one ran. The integer forms and FP64_FIXED are deterministic, so
the bar is byte equality.
"""

import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")

# (rows, features, depth, trees, form)
JOBS = [
    (40_000, 16, 6, 12, "int16"),
    (60_000, 12, 5, 12, "fp64_fixed"),
    (25_000, 20, 7, 10, "int8"),
    (1_100_000, 8, 4, 6, "int16"),
]


def _data(rows, feats, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats), dtype=np.float32)
    X[rng.random((rows, feats)) < 0.02] = np.nan
    y = np.sin(np.nan_to_num(X[:, 0])) + np.nan_to_num(X[:, 1]) * np.nan_to_num(X[:, 2]) + 0.1 * rng.standard_normal(rows)
    return X, y


_DATA = {}


def _fit(job):
    rows, feats, depth, trees, form = job
    X, y = _DATA[job]
    s = N._session()  # this thread's session
    ds = N._quantize_dense_device(s, X, 256)
    p = N._TrainingParameters(trees, 0.1, 256, depth, 20, 1.0, 1.0, precision=form, rounding_seed=11)
    m = N._train_binned(s, ds, y, None, p)
    return m.serialize(), m.fit_stats()


def test_two_sessions_match_sequential():
    for i, job in enumerate(JOBS):
        _DATA[job] = _data(job[0], job[1], 900 + i)
    print(f"concurrent sessions: {len(JOBS)} fits, sequential first", flush=True)
    sequential = [_fit(job) for job in JOBS]
    # Run the fits twice over on two threads, so every fit overlaps another one: thread 0
    # plans on the device (planner 2), thread 1 on the host (planner 1).
    print("concurrent sessions: 2 threads", flush=True)
    barrier = threading.Barrier(2)

    def worker(args):
        jobs, planner = args
        N._session().planner = planner
        try:
            barrier.wait(timeout=60)
            return [_fit(j) for j in jobs]
        finally:
            N._session().planner = 0

    with ThreadPoolExecutor(max_workers=2) as pool:
        device_run, host_run = pool.map(worker, [(JOBS, 2), (JOBS[::-1], 1)])
    host_run = host_run[::-1]
    for run in (device_run, host_run):
        for job, (blob, stats), (ref_blob, _) in zip(JOBS, run, sequential):
            assert blob == ref_blob, f"model bytes differ from the sequential fit for {job}"
            assert stats["trees"] == job[3]
    assert [st["device_planned_trees"] for _, st in device_run] == [job[3] for job in JOBS]
    assert [st["device_planned_trees"] for _, st in host_run] == [0] * len(JOBS)


def test_model_grid_on_concurrent_sessions_matches_sequential():
    # The 48-model grid workload in miniature: the grid's data formula, row counts of two of
    # its shape classes, its depths and learning rates, fitted through fit_concurrently on four
    # sessions and one after another. Every model's bytes must be the same.
    from mpsboost_amd import estimators as E

    def data(rows, seed):
        rng = np.random.default_rng(seed)
        X = rng.standard_normal((rows, 32), dtype=np.float32)
        y = np.sin(X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3] ** 2 + rng.normal(0.0, 0.1, rows)
        return X, y

    sets = {n: data(n, 5000 + n // 1000) for n in (50_000, 200_000)}

    def jobs():
        return [(E.GradientBoostingRegressor(n_estimators=15, max_depth=d, learning_rate=lr, min_samples_leaf=20,
                                             precision="int16"), *sets[n])
                for n in sets for d in (3, 6, 9) for lr in (0.03, 0.1)]

    print("model grid: 12 fits sequential, then on 4 sessions", flush=True)
    seq = E.fit_concurrently(jobs(), n_jobs=1)
    conc = E.fit_concurrently(jobs(), n_jobs=4)
    for a, b in zip(seq, conc):
        assert a.model_.serialize() == b.model_.serialize()


def test_random_forest_on_concurrent_sessions_matches_sequential():
    # RandomForestClassifier(n_jobs > 1) grows its trees on concurrent sessions
    # (src/mpsboost/estimators/forest.py:136-144 fans them out the same way): the trees and
    # probabilities must equal the sequential forest's.
    from mpsboost_amd import estimators as E

    rng = np.random.default_rng(77)
    X = rng.standard_normal((20_000, 12), dtype=np.float32)
    y = (X[:, 0] + 0.5 * X[:, 1] * X[:, 2] + 0.3 * rng.standard_normal(20_000) > 0).astype(np.float64)
    kw = dict(n_estimators=8, max_depth=8, random_state=5, precision="int16")
    print("random forest: 8 trees sequential, then on 4 sessions", flush=True)
    one = E.RandomForestClassifier(n_jobs=1, **kw).fit(X, y)
    four = E.RandomForestClassifier(n_jobs=4, **kw).fit(X, y)
    for a, b in zip(one.estimators_, four.estimators_):
        assert a.model_.serialize() == b.model_.serialize()
    assert np.array_equal(one.predict_proba(X), four.predict_proba(X))
