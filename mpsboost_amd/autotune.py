"""Autotuner for the histogram launch shape.

The histogram kernel's shape has four knobs (``mpsb_tuning`` in include/mpsboost_hip.h): threads
per block, 32-bit code words loaded per thread, the LDS bytes of a block's feature tile (how the
cells are split into tiles) and the number of blocks per level launch (rows per block). The best
values depend on the dataset's shape class: rows, features, code width and the histogram form.

``tune`` times short fits of the real trainer for each candidate on the device and keeps the
fastest; results are cached per shape class (in memory and, when a path is given, in a JSON
file). ``enable`` installs a hook so every later fit of the package applies the cached shape of
its class, tuning unseen classes on first use when ``tune_missing`` is set.
"""

from __future__ import annotations

import itertools
import json
import math
import os
import threading
import time
from typing import Iterable

import numpy as np

from . import _native

DEFAULT = {"lds_budget": 65544, "target_blocks": 1024, "min_slab_rows": 2048, "block": 1024, "words_per_thread": 4}

# Candidate grid (48 shapes): 256- or 1024-thread blocks, 2 or 4 words per thread, 32 or 64 KB
# tiles (four or two blocks per CU), 512 / 1024 / 2048 blocks per launch and 512- or 2048-row
# minimum slabs (the knob that matters for small row counts).
CANDIDATES = [
    {"block": b, "words_per_thread": w, "lds_budget": l, "target_blocks": t, "min_slab_rows": m}
    for b, w, l, t, m in itertools.product((256, 1024), (2, 4), (32776, 65544), (512, 1024, 2048), (512, 2048))
]

_cache: dict[str, dict[str, int]] = {}
_lock = threading.Lock()
_path: str | None = None


def shape_class(dataset: _native._BinnedDataset, precision: str, weighted: bool = False) -> str:
    """Rows to the nearest power of two, features to a multiple of 16, code width, form."""

    rows = 1 << max(0, round(math.log2(max(1, dataset.rows))))
    feats = 16 * max(1, math.ceil(dataset.features / 16))
    return f"r{rows}-f{feats}-c{dataset.code_bytes}-{precision}{'-w' if weighted else ''}"


def _load(path: str) -> None:
    if os.path.exists(path):
        with open(path) as fh:
            data = json.load(fh)
        with _lock:
            for key, value in data.items():
                _cache.setdefault(key, {k: int(v) for k, v in value.items()})


def _store(path: str) -> None:
    with _lock:
        data = dict(_cache)
    tmp = path + ".tmp"
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def lookup(key: str) -> dict[str, int] | None:
    with _lock:
        return None if key not in _cache else dict(_cache[key])


def _time_fit(session, dataset, labels, weights, parameters, trees: int, repeats: int) -> float:
    p = _native._TrainingParameters(
        trees, parameters.learning_rate, parameters.max_bins, parameters.max_depth, parameters.min_samples_leaf,
        parameters.min_child_weight, parameters.reg_lambda, reg_alpha=parameters.reg_alpha,
        max_delta_step=parameters.max_delta_step, gamma=parameters.gamma,
        min_gain_to_split=parameters.min_gain_to_split, objective=parameters.objective,
        split_strategy=parameters.split_strategy, precision=parameters.precision,
        rounding_seed=parameters.rounding_seed, objective_alpha=parameters.objective_alpha,
        tweedie_variance_power=parameters.tweedie_variance_power)
    best = math.inf
    for _ in range(repeats):
        t0 = time.perf_counter()
        _native._train_binned(session, dataset, labels, weights, p)
        best = min(best, time.perf_counter() - t0)
    return best


def tune(session: _native._Session, dataset: _native._BinnedDataset, labels, parameters: _native._TrainingParameters,
         weights=None, candidates: Iterable[dict[str, int]] | None = None, trees: int = 3, repeats: int = 2,
         path: str | None = None) -> dict[str, int]:
    """Time `trees`-tree fits (best of `repeats`) for each candidate shape on `dataset` and
    cache the fastest for its shape class. Returns the chosen shape, with the timings of all
    candidates under ``"_seconds"``."""

    y = np.ascontiguousarray(labels, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    before = session.tuning
    results = []
    try:
        session.tuning = DEFAULT
        _time_fit(session, dataset, y, w, parameters, 1, 1)  # warm-up: upload, allocations
        for cand in (candidates or CANDIDATES):
            try:
                session.tuning = cand
            except _native.BackendError:
                continue  # e.g. an LDS budget above this device's maximum
            results.append((_time_fit(session, dataset, y, w, parameters, trees, repeats), dict(cand)))
    finally:
        session.tuning = before
    seconds, best = min(results, key=lambda item: item[0])
    key = shape_class(dataset, parameters.precision, w is not None)
    with _lock:
        _cache[key] = best
    if path or _path:
        _store(path or _path)
    out = dict(best)
    out["_seconds"] = [(s, c) for s, c in results]
    return out


def synthetic_labels(objective: str, rows: int, seed: int = 0) -> np.ndarray:
    """Labels valid for `objective`, for timing only."""

    rng = np.random.default_rng(seed)
    if objective == "binary_logistic":
        return (rng.random(rows) < 0.5).astype(np.float64)
    if objective in ("poisson", "tweedie"):
        return rng.poisson(1.0, rows).astype(np.float64)
    return rng.standard_normal(rows)


def enable(path: str | None = None, tune_missing: bool = False, trees: int = 3) -> None:
    """Apply cached shapes to every later fit; with `tune_missing`, tune a shape class the first
    time a fit meets it (that fit then pays for the search)."""

    global _path
    _path = path
    if path:
        _load(path)

    def hook(session, dataset, parameters):
        key = shape_class(dataset, parameters.precision)
        shape = lookup(key)
        if shape is None and tune_missing:
            shape = tune(session, dataset, synthetic_labels(parameters.objective, dataset.rows), parameters,
                         trees=trees, repeats=1)
        session.tuning = {k: v for k, v in (shape or DEFAULT).items() if not k.startswith("_")}

    _native._tuning_hook = hook


def disable() -> None:
    _native._tuning_hook = None
