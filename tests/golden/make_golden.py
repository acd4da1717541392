"""Writes tests/golden/reference_vectors.npz: inputs and the reference's own outputs.

The outputs come from the reference's C++ sources compiled unchanged into oracle/_ref/libref.so
(oracle/Makefile, target `ref`): QuantizeDense, CpuReferenceBackend::BuildHistograms,
TrainRegressionModel and TrainMulticlassSoftmaxModel with RegressionTree/Model prediction.
Run from the repository root in a checkout next to the reference sources:

    python tests/golden/make_golden.py

The file holds plain arrays only (no pickled objects); tests load it with numpy's default
allow_pickle=False.
"""

from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")

# Tree cases: (name, max_bins, n_estimators, learning_rate, tree parameters).
TREE_CASES = [
    ("plain", 32, 3, 0.3, dict(max_depth=5, min_samples_leaf=3)),
    ("regularised", 32, 3, 0.3, dict(max_depth=4, reg_lambda=0.5, reg_alpha=0.2, max_delta_step=0.6, gamma=0.05,
                                      min_child_weight=2.0)),
    ("constrained", 16, 3, 0.5, dict(max_depth=4, monotonic_constraints=[1, 0, -1, 0, 0, 0],
                                      interaction_constraints=[[0, 1, 2], [3, 4, 5]], min_samples_leaf=4)),
    ("random_threshold", 32, 2, 0.3, dict(max_depth=3, split_strategy="random_threshold", random_seed=11)),
]


def inputs(rows=1800, feats=6, seed=20240):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    X[:, 1] = rng.integers(0, 6, rows)
    X[:, 3] = np.round(X[:, 3], 1)
    X[rng.random(X.shape) < 0.05] = np.nan
    y = (np.sin(np.nan_to_num(X[:, 0])) + np.nan_to_num(X[:, 2]) * np.nan_to_num(X[:, 4])
         + 0.3 * np.nan_to_num(X[:, 1]) + 0.05 * rng.standard_normal(rows))
    labels3 = rng.integers(0, 3, rows).astype(np.float64)
    labels3[np.nan_to_num(X[:, 0]) > 0.8] = 2.0
    rows_sel = np.sort(rng.choice(rows, rows // 3, replace=False)).astype(np.uint64)
    return X, y, labels3, rows_sel


def main() -> int:
    ref = O.reference()
    if ref is None:
        print("oracle/_ref/libref.so is not built (reference sources absent)", file=sys.stderr)
        return 1
    X, y, labels3, rows_sel = inputs()
    out = {"X": X, "y": y, "labels3": labels3, "rows_sel": rows_sel}
    for mb in (16, 32, 300):
        ds = ref.quantize(X, mb)
        out[f"q{mb}_bins"] = ds.bins.astype(np.uint16)
        out[f"q{mb}_missing"] = ds.missing.astype(np.uint8)
        out[f"q{mb}_bin_counts"] = ds.bin_counts
        out[f"q{mb}_boundaries"] = np.concatenate([np.asarray(b, dtype=np.float32) for b in ds.boundaries] or
                                                  [np.zeros(0, np.float32)])
    ds32 = ref.quantize(X, 32)
    pairs = ref.squared_error(y, np.zeros_like(y))
    out["pairs"] = pairs
    c, g, h = ref.histograms(ds32, pairs, rows_sel)
    out["hist32_count"], out["hist32_g"], out["hist32_h"] = c, g, h
    for name, mb, n, lr, kw in TREE_CASES:
        ds = ref.quantize(X, mb)
        m = ref.train_regressor(ds, y, n_estimators=n, learning_rate=lr, **kw)
        out[f"tree_{name}_nodes"] = np.concatenate(m.trees)
        out[f"tree_{name}_sizes"] = np.asarray([t.size for t in m.trees], dtype=np.uint32)
        out[f"tree_{name}_pred"] = m.predict(ds)
    m = ref.train_multiclass(ds32, labels3, 3, n_estimators=2, learning_rate=0.3, max_depth=3, min_samples_leaf=5)
    out["multi_nodes"] = np.concatenate(m.trees)
    out["multi_sizes"] = np.asarray([t.size for t in m.trees], dtype=np.uint32)
    out["multi_pred"] = m.predict(ds32)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT):,} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
