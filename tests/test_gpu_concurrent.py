"""Concurrent device sessions: regression fits on two threads at once, each thread on its own
session (one HIP stream each), give model bytes identical to the same fits run one after
another. The reference fans independent fits out over threads the same way
(src/mpsboost/estimators/forest.py:136-144, n_jobs).

Fits under 1M rows that start while another fit is live are planned on the host: every level
ends in a host wait for the device-published split candidates (the bounded wait of
mpsboost_amd/csrc/host_wait.hpp). The test checks that path ran (device_planned_trees == 0 for
those fits) and that a fit of at least 1M rows kept the device planner. The integer forms and
FP64_FIXED are deterministic, so the bar is byte equality.
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
    # Run the fits twice over on two threads, so every fit overlaps another one.
    print("concurrent sessions: 2 threads", flush=True)
    barrier = threading.Barrier(2)

    def worker(jobs):
        barrier.wait(timeout=60)
        return [_fit(j) for j in jobs]

    with ThreadPoolExecutor(max_workers=2) as pool:
        halves = list(pool.map(worker, [JOBS, JOBS[::-1]]))
    concurrent = [halves[0], halves[1][::-1]]
    for run in concurrent:
        for job, (blob, stats), (ref_blob, _) in zip(JOBS, run, sequential):
            assert blob == ref_blob, f"model bytes differ from the sequential fit for {job}"
            assert stats["trees"] == job[3]
    # Which planner ran: the first thread's 1.1M-row fit kept device planning; small fits that
    # overlapped a live fit were host-planned (level round trips through the bounded wait).
    host_planned = [stats["device_planned_trees"] for run in concurrent for (_, stats), job in zip(run, JOBS)
                    if job[0] < 1_000_000]
    assert any(t == 0 for t in host_planned), "no small fit ran the host-planned concurrent path"
    big = [stats["device_planned_trees"] for run in concurrent for (_, stats), job in zip(run, JOBS)
           if job[0] >= 1_000_000]
    assert all(t == JOBS[-1][3] for t in big)
