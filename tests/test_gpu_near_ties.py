"""Near-tie resolution of the stochastic-rounding forms (device-planned trees).

The INT16 and INT8 forms round each row's gradient stochastically, so a level's gains carry
rounding noise with a known bound. A node whose best candidate that noise could have swapped
(another candidate's gain plus its bound reaches the best's gain minus its bound, or the best
split's acceptance is in doubt) is scanned again over its contender features from exact 64-bit
fixed-point histograms of the round's FP64 gradients: the arithmetic of the FP64_FIXED form.
With leaf renewal, which solves every leaf from exact fixed-point sums of the same gradients,
every tree of a quantised fit then equals the FP64_FIXED fit's tree: same features,
thresholds, default directions, children and leaf values (the stored gains of nodes that were
not ambiguous stay the form's own).

The host planner and the device planner resolve the same nodes the same way.

The data make near ties on purpose: a few features are near-copies of the informative ones,
so their best gains differ by far less than the rounding noise.

The reference (src/core/objective.cpp:303-318) quantises gradients with no such resolution;
the bar here is the FP64_FIXED form of this backend, which the oracle pins separately
(test_gpu_configs.py, test_gpu_parity.py).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")

ROWS, FEATURES, DEPTH, TREES = 200_000, 16, 5, 40
STRUCTURE = ("feature_index", "threshold_bin", "left_child", "right_child", "default_left", "flags", "leaf_value")


def _data(seed=31):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((ROWS, FEATURES), dtype=np.float32)
    # Near-copies of the informative features: gains within the rounding noise of each other.
    X[:, 8:11] = X[:, 0:3] + np.float32(1e-3) * rng.standard_normal((ROWS, 3), dtype=np.float32)
    X[rng.random((ROWS, FEATURES)) < 0.01] = np.nan
    Z = np.nan_to_num(X)
    y = np.sin(Z[:, 0]) + Z[:, 1] * Z[:, 2] + 0.5 * Z[:, 3] + 0.3 * rng.standard_normal(ROWS)
    return X, y


@pytest.fixture(scope="module")
def fitted():
    s = N._session()
    X, y = _data()
    ds = N._quantize_dense_device(s, X, 256)
    previous = s.planner
    try:
        def fit(form, near_ties="resolve", planner=2, seed=5):
            s.planner = planner  # 2: levels planned on the device, 1: on the host
            p = N._TrainingParameters(TREES, 0.1, 256, DEPTH, 1, 0.0, 1.0, precision=form, rounding_seed=seed,
                                      near_ties=near_ties)
            m = N._train_binned(s, ds, y, None, p)
            return m.trees, m.fit_stats()

        out = {"fp64_fixed": fit("fp64_fixed")}
        for form in ("int16", "int8"):
            print(f"near ties: {form} resolve", flush=True)
            out[form] = fit(form)
            out[form + "_keep"] = fit(form, "keep")
        out["int16_host"] = fit("int16", planner=1)
        out["int16_auto"] = fit("int16", "auto")
        out["int8_auto"] = fit("int8", "auto")
        return out
    finally:
        s.planner = previous


def _first_difference(trees, ref):
    for t, (a, b) in enumerate(zip(trees, ref)):
        if a.size != b.size:
            return t, "node count"
        for k in STRUCTURE:
            if not np.array_equal(a[k], b[k]):
                return t, k
    return None


@pytest.mark.parametrize("form", ["int16", "int8", "int16_host", "int16_auto"])
def test_resolved_fit_equals_fp64_fixed(fitted, form):
    ref, ref_stats = fitted["fp64_fixed"]
    trees, st = fitted[form]
    assert ref_stats["ambiguous_nodes"] == 0  # exact form: no rounding noise, nothing to resolve
    assert st["device_planned_trees"] == (0 if form == "int16_host" else TREES)
    assert st["renewed_trees"] == TREES
    assert st["ambiguous_nodes"] > 0, "the near-copy features made no ambiguous node"
    # Every level holds at most 2^(depth - 1) nodes, fewer than the 64 resolved per level.
    assert st["resolved_nodes"] == st["ambiguous_nodes"]
    assert len(trees) == len(ref) == TREES
    diff = _first_difference(trees, ref)
    assert diff is None, f"{form}: tree {diff[0]} differs from the FP64_FIXED fit in {diff[1]}"


@pytest.mark.parametrize("form", ["int16_keep", "int8_keep", "int8_auto"])
def test_kept_ties_skip_resolution(fitted, form):
    """"keep" takes the rounded data's choice at near-ties; "auto" keeps them in int8, whose
    noise makes most nodes ambiguous."""
    trees, st = fitted[form]
    assert st["ambiguous_nodes"] == 0 and st["resolved_nodes"] == 0
    assert len(trees) == TREES
