"""The benchmark workloads at their stated sizes, on the device, against the CPU oracle.

* 100,000 x 28 regression (N(0,1) features from seed 1001, y = sin x0 + x1 x2 + 0.5 x3^2 +
  N(0, 0.1^2) from seed 1002; 256 bins; 100 trees, depth 6, lr 0.1, min_child_weight 1,
  lambda 1, min_samples_leaf 20):
  - FP64 form: every tree, gain, leaf and prediction bit-identical to the oracle;
  - FP32 and FP64_FIXED forms, tree by tree on the device's own predictions (the oracle grows
    tree i from the FP64 gradients of the device's predictions after i - 1 trees, so no
    trajectory drift enters): identical structure and row partitions, except at a node where
    the oracle's best candidate and the device's differ in exact gain by no more than the
    derived error bound of the form's sums (below); leaves within 1e-5 relative (to the leaf, or
    where its gradient sum cancels, to sum |g| / (H + lambda) of its rows, the scale a sum's
    rounding is relative to). The number of identical trees is asserted and printed.
* 10,000,000 x 128 regression (the bench workload, 20 trees, depth 8): the quantised forms and
  FP64_FIXED bit-identical across two runs and two launch shapes; the histogram row count of
  the fit equal to the one implied by routing every row through the trees on the CPU (the
  root plus the smaller child of every split above the last level); device predictions equal
  to the CPU's tree-order sums; FP64_FIXED trees on the first 1,000,000 rows equal to the
  oracle's up to near-ties.
* 4,000,000 x 64 softmax, K = 10 (one round of ten class trees, depth 7, int16): each class tree
  equal to the oracle's grown from the same stochastically rounded gradients (replayed from the
  device's counter-based stream) up to near-ties.

Error bound of a candidate's gain (FP32 and FP64_FIXED forms). Against the exact FP64 sums of
the oracle's gradients, a device sum over a set of rows errs by at most
  |sum of the rows' gradient differences| (exact: FP32 gradients are float(p) - float(y)) plus,
  per histogram cell summed in float chains of at most m values, gamma_m * sum |value|
  (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., 4.2: gamma_m = m u /
  (1 - m u), u = 2^-24; m = min(the cell's rows, one slab's 2048 rows + 16 replicas));
or, in the fixed-point form, rows * 2^-(S+1) (each row rounded to a multiple of 2^-S, integer
sums exact). A sibling's cells are parent - built, so they carry both histograms' summation
bounds; a node's statistics come from its ancestors' prefix scans, so their bounds add up the
path. With eps_L, eps_P (parent) and eps_R = eps_P + eps_L, the gain
0.5 (G_L^2/(H_L+l) + G_R^2/(H_R+l) - G_P^2/(H_P+l)) moves by at most
sum_k (|G_k| eps_k + eps_k^2 / 2) / (H_k + l) (Hessians are exact counts here). Two candidates
whose exact gains differ by less than the sum of their bounds may be ranked either way. On
this data the bound is about 4e-4 of the root's gain: a root threshold one bin off is caught.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")

STRUCT = ["feature_index", "threshold_bin", "left_child", "right_child", "default_left", "flags"]
KEY = ["flags", "feature_index", "threshold_bin", "default_left"]
U32 = 2.0 ** -24
THREADS = min(16, len(os.sched_getaffinity(0)))


@pytest.fixture(scope="module")
def oracle():
    o = O.oracle()
    o.set_threads(THREADS)
    yield o
    o.set_threads(1)


@pytest.fixture(scope="module")
def backend():
    return N._HipBackend()


def _params(n, lr, depth, precision, msl=20, mcw=1.0, **kw):
    return N._TrainingParameters(n, lr, 256, depth, msl, mcw, 1.0, precision=precision, **kw)


# ------------------------------------------------------------------ 100k x 28 -------
C0 = dict(rows=100_000, features=28, trees=100, depth=6, lr=0.1, msl=20, mcw=1.0)


def _config0_data():
    rng = np.random.default_rng(1001)
    X = np.stack([rng.standard_normal(C0["rows"]) for _ in range(C0["features"])], axis=1).astype(np.float32)
    x = X.astype(np.float64)
    y = np.sin(x[:, 0]) + x[:, 1] * x[:, 2] + 0.5 * x[:, 3] ** 2 + np.random.default_rng(1002).normal(0, 0.1, C0["rows"])
    return X, y


@pytest.fixture(scope="module")
def config0(oracle):
    X, y = _config0_data()
    return X, y, oracle.quantize(X, 256), N._quantize_dense(X, 256)


def _route(ods, nodes):
    """Node index of every row at every depth (-1 past its leaf) and the leaf of every row."""

    rows = np.arange(ods.rows)
    node = np.zeros(ods.rows, dtype=np.int64)
    out = [node.copy()]
    for _ in range(64):
        n = nodes[node]
        leaf = (n["flags"] & 1) == 1
        if leaf.all():
            break
        f = n["feature_index"].astype(np.int64)
        left = np.where(ods.missing[f, rows], n["default_left"] == 1, ods.bins[f, rows] <= n["threshold_bin"])
        node = np.where(leaf, node, np.where(left, n["left_child"], n["right_child"]).astype(np.int64))
        out.append(np.where(leaf, -1, node))
    return np.stack(out, axis=1), node


def test_config0_fp64_bit_identical(backend, oracle, config0):
    X, y, ods, ds = config0
    print("100k x 28: fp64 form, 100 trees", flush=True)
    model = N._train_binned(backend.session, ds, y, None,
                            _params(C0["trees"], C0["lr"], C0["depth"], "fp64", C0["msl"], C0["mcw"]))
    ref = oracle.train_regressor(ods, y, n_estimators=C0["trees"], learning_rate=C0["lr"], max_depth=C0["depth"],
                                 min_samples_leaf=C0["msl"], min_child_weight=C0["mcw"])
    assert model.tree_count == ref.tree_count == C0["trees"]
    for i, (a, b) in enumerate(zip(model.trees, ref.trees)):
        for k in STRUCT + ["leaf_value", "gain"]:
            assert np.array_equal(a[k], b[k]), f"tree {i}: {k}"
    assert np.array_equal(model.predict_binned(backend.session, ds), ref.predict(ods))


def _gain(gl, hl, gp, hp, lam):
    gr, hr = gp - gl, hp - hl
    return 0.5 * (gl * gl / (hl + lam) + gr * gr / (hr + lam) - gp * gp / (hp + lam))


class _NodeBound:
    """Error bounds of the device's sums at node k of a tree grown like `tree` (the device's
    tree equals it up to k). Per row: d (signed error of the row's gradient in the form's
    arithmetic, exact), a (an absolute bound on further per-row error), w (the row's |value|,
    weight of the summation term); chain: the longest accumulation chain of one cell (0 for
    exact integer sums).

    A cell summed from c rows in chains of at most `chain` values errs by at most
    gamma_min(c, chain) * sum |w| over its rows. Built nodes sum their own rows; a sibling is
    parent - built, so its cells carry the summation errors of its parent's histogram and of
    the built child's (`hists`). The node's statistics come from its parent's prefix scan over
    the split feature (and, for a right child, the parent's statistics), chained up to the
    root, whose statistics are its histogram totals."""

    def __init__(self, ods, tree, paths, k, d, a, w, chain):
        self.ods, self.d, self.a, self.w, self.chain = ods, d, a, w, chain
        counts = np.bincount(paths[paths >= 0], minlength=tree.size)
        parent = np.full(tree.size, -1)
        for i in np.flatnonzero((tree["flags"] & 1) == 0):
            parent[tree["left_child"][i]] = parent[tree["right_child"][i]] = i
        self.member = {}
        path = [k]
        while parent[path[-1]] >= 0:
            path.append(int(parent[path[-1]]))
        root = path[-1]
        hists = {root: [root]}
        for n in reversed(path[:-1]):
            p = int(parent[n])
            left = tree["left_child"][p] == n
            sib = int(tree["right_child"][p] if left else tree["left_child"][p])
            built = counts[n] < counts[sib] or (counts[n] == counts[sib] and left)
            hists[n] = [n] if built else hists[p] + [sib]
        self.paths, self.hists_k = paths, hists[k]
        self.rows = self._rows(k)
        self._cells = {}
        eps = abs(float(d[self.rows].sum())) + float(a[self.rows].sum())
        eps += max(self._total(root, f) for f in range(ods.features))  # the root's statistics
        for anc in path[1:]:  # each ancestor's prefix scan over its split feature
            f = int(tree["feature_index"][anc])
            eps += sum(self._total(h, f) for h in hists[anc])
        self.eps_p = eps

    def _rows(self, n):
        if n not in self.member:
            self.member[n] = (self.paths == n).any(axis=1)
        return self.member[n]

    def _cell_terms(self, h, f):
        """Summation bound of every cell of feature f in histogram h: bins, then missing."""

        if (h, f) not in self._cells:
            r = self._rows(h)
            b = self.ods.bins[f][r]
            m = self.ods.missing[f][r]
            w = self.w[r]
            nb = int(self.ods.bin_counts[f])
            cnt = np.append(np.bincount(b[~m], minlength=nb), m.sum()).astype(np.float64)
            sw = np.append(np.bincount(b[~m], weights=w[~m], minlength=nb), w[m].sum())
            n = np.minimum(cnt, self.chain) * U32
            self._cells[(h, f)] = n / (1.0 - n) * sw
        return self._cells[(h, f)]

    def _total(self, h, f):
        return float(self._cell_terms(h, f).sum()) if self.chain else 0.0

    def _left_mask(self, f, t, dl):
        miss = self.ods.missing[f]
        return self.rows & ((~miss & (self.ods.bins[f] <= t)) | (miss & dl))

    def candidate(self, node, g, lam):
        """Exact gain of the node's split (0 for a leaf), the bound on the device's error in
        it, and the two child counts."""

        if node["flags"] & 1:
            return 0.0, 0.0, None
        f, t, dl = int(node["feature_index"]), int(node["threshold_bin"]), bool(node["default_left"])
        sel = self._left_mask(f, t, dl)
        gl, hl = float(g[sel].sum()), float(sel.sum())
        gp, hp = float(g[self.rows].sum()), float(self.rows.sum())
        eps_l = abs(float(self.d[sel].sum())) + float(self.a[sel].sum())
        if self.chain:
            for h in self.hists_k:
                terms = self._cell_terms(h, f)
                eps_l += float(terms[: t + 1].sum()) + (float(terms[-1]) if dl else 0.0)
        eps_r = self.eps_p + eps_l
        bound = 0.0
        scale = 0.0
        for gk, hk, ek in ((gl, hl, eps_l), (gp - gl, hp - hl, eps_r), (gp, hp, self.eps_p)):
            bound += (abs(gk) * ek + ek * ek / 2.0) / (hk + lam)
            scale += gk * gk / (hk + lam)
        return _gain(gl, hl, gp, hp, lam), bound + 1e-12 * scale, (hl, hp - hl)


def _teacher_forced(model, oracle, ods, y, lr, depth, msl, mcw, row_errors, where):
    """Tree i of the device against the oracle's tree grown from the device's predictions
    after i - 1 trees. row_errors(p, y, g) -> (d, a, w, chain) of _NodeBound. Returns
    (identical trees, near-tie trees, largest relative leaf difference, final predictions)."""

    lam = 1.0
    p = np.full(ods.rows, model.base_score)
    identical = near = 0
    worst_leaf = 0.0
    for i, tree in enumerate(model.trees):
        pairs = oracle.squared_error(y, p)
        ref = oracle.train_tree(ods, pairs, max_depth=depth, min_samples_leaf=msl, min_child_weight=mcw,
                                reg_lambda=lam)
        n = min(tree.size, ref.size)
        diff = np.flatnonzero(np.any(np.stack([tree[k][:n] != ref[k][:n] for k in KEY], axis=1), axis=1))
        if tree.size == ref.size and diff.size == 0 and all(np.array_equal(tree[k], ref[k]) for k in STRUCT):
            identical += 1
            # Relative to the leaf's magnitude, or, where its gradient sum cancels, to the
            # scale of the gradients it sums: sum |g| / (H + lambda) over the leaf's rows.
            _, leaf_of = _route(ods, ref)
            size = np.bincount(leaf_of, minlength=ref.size)
            spread = np.bincount(leaf_of, weights=np.abs(pairs[:, 0]), minlength=ref.size) / (size + lam)
            leaf = (ref["flags"] & 1) == 1
            rel = np.abs(tree["leaf_value"][leaf] - ref["leaf_value"][leaf]) / np.maximum(
                np.abs(ref["leaf_value"][leaf]), spread[leaf])
            worst_leaf = max(worst_leaf, float(rel.max()))
            assert rel.max() <= 1e-5, f"{where} tree {i}: leaf values differ by {rel.max():.3g} relative"
        else:
            k = int(diff[0]) if diff.size else n
            paths, _ = _route(ods, ref)
            g = pairs[:, 0]
            nb = _NodeBound(ods, ref, paths, k, *row_errors(p, y, g))
            go, bo, _ = nb.candidate(ref[k], g, lam)
            gd, bd, hd = nb.candidate(tree[k], g, lam)
            if not ref[k]["flags"] & 1:
                assert go == pytest.approx(float(ref[k]["gain"]), rel=1e-9, abs=1e-9), "exact gain formula"
            if hd is not None:
                assert min(hd) >= msl, f"{where} tree {i} node {k}: device split violates min_samples_leaf"
            assert go - gd <= bo + bd, (f"{where} tree {i} node {k}: gains {go!r} (oracle) vs {gd!r} (device) differ "
                                        f"by {go - gd:.3g}, beyond the rounding bound {bo + bd:.3g}")
            near += 1
            print(f"{where} tree {i} node {k}: near-tie, gain gap {go - gd:.3g} within bound {bo + bd:.3g}",
                  flush=True)
        _, leaf_of = _route(ods, tree)
        p = p + lr * tree["leaf_value"][leaf_of]
    return identical, near, worst_leaf, p


# One float accumulation chain of the FP32 form adds at most one slab's rows of a cell
# (slabs of at most 2048 rows below 2M rows a level at the default launch shape) plus the 16
# row-slot replicas of a narrow feature at the flush.
FP32_CHAIN = 2048 + 16


def _fp32_row_errors(p, y, g):
    g32 = (p.astype(np.float32) - y.astype(np.float32)).astype(np.float64)
    return g32 - g, np.zeros_like(g), np.abs(g32), FP32_CHAIN


def _fixed_row_errors(p, y, g):
    """Fixed point: each row rounded to a multiple of 2^-S, S = 62 - exponent(max|g| * rows)."""

    _, e = np.frexp(np.max(np.abs(g)) * g.size)
    half_ulp = 2.0 ** -(62 - int(e) + 1)
    return np.zeros_like(g), np.full_like(g, half_ulp), np.zeros_like(g), 0


@pytest.mark.parametrize("form", ["fp32", "fp64_fixed"])
def test_config0_structure_within_derived_bound(backend, oracle, config0, form):
    X, y, ods, ds = config0
    print(f"100k x 28: {form} form, 100 trees, tree by tree against the oracle", flush=True)
    model = N._train_binned(backend.session, ds, y, None,
                            _params(C0["trees"], C0["lr"], C0["depth"], form, C0["msl"], C0["mcw"]))
    assert model.tree_count == C0["trees"]
    errors = _fp32_row_errors if form == "fp32" else _fixed_row_errors
    identical, near, worst, p = _teacher_forced(model, oracle, ods, y, C0["lr"], C0["depth"], C0["msl"], C0["mcw"],
                                                errors, form)
    print(f"100k x 28 {form}: {identical} of {C0['trees']} trees identical to the oracle, {near} part at a "
          f"near-tie within the bound; largest leaf difference {worst:.2e} relative", flush=True)
    assert identical + near == C0["trees"]
    assert identical >= 90, f"only {identical} trees identical"
    # The device's own predictions are the same tree-order FP64 sums.
    assert np.array_equal(model.predict_binned(backend.session, ds), p)


# -------------------------------------------------------------- 10M x 128 ---------
C1 = dict(rows=10_000_000, trees=20, depth=8, lr=0.05, msl=1, mcw=0.0)
ALT_SHAPE = {"block": 256, "words_per_thread": 2, "lds_budget": 32768, "target_blocks": 512, "min_slab_rows": 512}


@pytest.fixture(scope="module")
def config1(oracle, backend):
    from bench import make_data

    print("10M x 128: generating the data", flush=True)
    X, y = make_data(C1["rows"], 2001)
    ds = N._quantize_dense_device(backend.session, X, 256)
    print("10M x 128: binned on the device", flush=True)
    return X, y, ds


def _fit(backend, ds, y, form, trees=C1["trees"], rows_params=C1, **kw):
    p = N._TrainingParameters(trees, rows_params["lr"], 256, rows_params["depth"], rows_params["msl"],
                              rows_params["mcw"], 1.0, precision=form, rounding_seed=7, **kw)
    return N._train_binned(backend.session, ds, y, None, p)


def _built_rows(oracle, ods, tree, depth):
    """Rows histogrammed for one tree: the root plus the smaller child (the left one on ties)
    of every split whose children lie above the depth limit, from the rows routed on the CPU."""

    ids = tree.copy()
    ids["leaf_value"] = np.arange(tree.size, dtype=np.float64)  # tree_predict returns the leaf's index
    leaf = oracle.tree_predict(ods, ids, ods.features).astype(np.int64)
    counts = np.bincount(leaf, minlength=tree.size).astype(np.int64)
    node_depth = np.zeros(tree.size, dtype=np.int64)
    internal = np.flatnonzero((tree["flags"] & 1) == 0)
    for i in internal:  # BFS order: parents come before their children
        node_depth[tree["left_child"][i]] = node_depth[tree["right_child"][i]] = node_depth[i] + 1
    for i in internal[::-1]:  # children before parents
        counts[i] = counts[tree["left_child"][i]] + counts[tree["right_child"][i]]
    built = counts[0]
    for i in internal:
        if node_depth[i] + 1 < depth:
            built += min(counts[tree["left_child"][i]], counts[tree["right_child"][i]])
    return int(built), counts


def test_config1_forms_deterministic_counts_and_predictions(backend, oracle, config1):
    X, y, ds = config1
    s = backend.session
    blobs = {}
    for form in ("int16", "int8", "fp64_fixed"):
        print(f"10M x 128: {form}, two runs and a second launch shape", flush=True)
        a = _fit(backend, ds, y, form).serialize()
        b = _fit(backend, ds, y, form).serialize()
        before = s.tuning
        s.tuning = ALT_SHAPE
        try:
            c = _fit(backend, ds, y, form).serialize()
        finally:
            s.tuning = before
        assert a == b, f"{form}: two runs differ"
        assert a == c, f"{form}: the launch shape changed the model"
        blobs[form] = a
    print("10M x 128: binning on the CPU oracle for routing", flush=True)
    ods = oracle.quantize(X, 256)
    for form in ("int16", "fp32"):
        model = _fit(backend, ds, y, form)
        st = model.fit_stats()
        p = np.full(C1["rows"], model.base_score)
        total = 0
        for tree in model.trees:
            built, counts = _built_rows(oracle, ods, tree, C1["depth"])
            assert counts[0] == C1["rows"]
            leaves = np.flatnonzero(tree["flags"] & 1)
            assert counts[leaves].min() >= C1["msl"]
            total += built
            p = p + C1["lr"] * oracle.tree_predict(ods, tree, ods.features)
        print(f"10M x 128 {form}: {total:,} histogram rows over {model.tree_count} trees", flush=True)
        assert st["histogram_rows"] == total, f"{form}: device counted {st['histogram_rows']}, routing gives {total}"
        assert np.array_equal(model.predict_binned(s, ds), p), f"{form}: predictions differ from tree-order sums"


def test_config1_fp64_fixed_prefix_matches_oracle(backend, oracle, config1):
    """The first 1,000,000 rows, binned on their own: FP64_FIXED trees against the oracle's,
    tree by tree on the device's predictions, within the fixed-point rounding bound."""

    X, y, _ = config1
    n = 1_000_000
    Xp, yp = np.ascontiguousarray(X[:n]), y[:n]
    ods = oracle.quantize(Xp, 256)
    ds = N._quantize_dense_device(backend.session, Xp, 256)
    assert np.array_equal(ds.bins, ods.bins) and np.array_equal(ds.missing, ods.missing)
    model = _fit(backend, ds, yp, "fp64_fixed", trees=3)
    print("1M-row prefix: fp64_fixed, 3 trees against the oracle", flush=True)
    identical, near, worst, p = _teacher_forced(model, oracle, ods, yp, C1["lr"], C1["depth"], C1["msl"], C1["mcw"],
                                                _fixed_row_errors, "1M prefix fp64_fixed")
    print(f"1M-row prefix: {identical} of 3 trees identical, {near} near-ties; leaves within {worst:.2e}", flush=True)
    assert identical + near == 3
    assert np.array_equal(model.predict_binned(backend.session, ds), p)


# ------------------------------------------------------------- softmax 4M x 64 -----
def _mix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return z ^ (z >> np.uint64(31))


def _round_channel(v, seed, ch, qmax):
    """The device's stochastic rounding of one channel (0: g, 1: h): multiplier qmax / max|v|,
    floor(v * m + u) with u from splitmix64 keyed by (round seed, 2 * row + channel)."""

    m = np.max(np.abs(v))
    i = np.arange(v.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        bits = _mix64(np.uint64(seed) ^ _mix64(i * np.uint64(2) + np.uint64(ch)))
    u = (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.clip(np.floor(v * (qmax / m) + u), -qmax, qmax) * (m / qmax)


def _leaf_ids(oracle, ods, tree):
    ids = tree.copy()
    ids["leaf_value"] = np.arange(tree.size, dtype=np.float64)
    return oracle.tree_predict(ods, ids, ods.features).astype(np.int64)


def _check_renewed_leaves(oracle, ods, tree, g, h, lam, where):
    """Leaf renewal: every leaf of the device's tree equals the oracle's leaf solve
    (LeafWeight) of the exact sums of the round's FP64 (g, h) over the leaf's rows (math.fsum:
    correctly rounded). The device's fixed-point sums err by at most rows * 2^-(S+1), far below
    the 1e-9 of sum |g| / (H + lambda) allowed here, and far below the difference to leaves
    from the quantised histogram (about 1e-5)."""

    import math

    leaf_of = _leaf_ids(oracle, ods, tree)
    order = np.argsort(leaf_of, kind="stable")
    bounds = np.concatenate([[0], np.cumsum(np.bincount(leaf_of, minlength=tree.size))])
    worst = 0.0
    for node in np.flatnonzero(tree["flags"] & 1):
        rows = order[bounds[node]: bounds[node + 1]]
        G, H = math.fsum(g[rows]), math.fsum(h[rows])
        expect = oracle.leaf_weight(G, H, lam)
        scale = max(abs(expect), math.fsum(np.abs(g[rows])) / (H + lam))
        gap = abs(tree["leaf_value"][node] - expect)
        worst = max(worst, gap / scale)
        assert gap <= 1e-9 * scale, f"{where} leaf {node}: {tree['leaf_value'][node]!r} vs the exact solve {expect!r}"
    return worst


def test_config3_softmax_round_matches_oracle_on_same_rounding(backend, oracle):
    """One round of K = 10 class trees at 4M x 64, int16: class tree c against the oracle's tree
    grown from the same rounded (g, h) of the softmax over the device's margins after the
    round's earlier class trees; equal up to near-ties (gains of the same rounded data within
    1e-9 relative: exact integer sums against row-order FP64 sums). Leaves are renewed: each
    equals the exact solve from the unrounded FP64 (g, h) of its rows."""

    from bench import make_softmax_data

    K, lr, seed, depth = 10, 0.1, 7, 7
    X, y = make_softmax_data(4_000_000)
    ods = oracle.quantize(X, 256)
    ds = N._quantize_dense_device(backend.session, X, 256)
    # The rounded data's own choice at near-ties, as the oracle makes it.
    p = N._TrainingParameters(1, lr, 256, depth, 20, 1.0, 1.0, precision="int16", rounding_seed=seed,
                              near_ties="keep")
    model = N._train_binned(backend.session, ds, y, None, p, class_count=K)
    assert model.tree_count == K
    margins = np.tile(model.base_scores, (y.size, 1))
    identical = near = 0
    worst = 0.0
    for t, tree in enumerate(model.trees):
        c = t % K
        e = np.exp(margins - margins.max(axis=1, keepdims=True))
        z = np.zeros(y.size)
        for k in range(K):  # class order, as the device sums
            z = z + e[:, k]
        prob = e[:, c] / z
        g, h = prob - (y == c), prob * (1.0 - prob)
        s = (seed + t * 0x9E3779B97F4A7C15) % (1 << 64)
        pairs = np.stack([_round_channel(g, s, 0, 32767.0), _round_channel(h, s, 1, 32767.0)], axis=1)
        ref = oracle.train_tree(ods, pairs, max_depth=depth, min_samples_leaf=20, min_child_weight=1.0, reg_lambda=1.0)
        n = min(tree.size, ref.size)
        diff = np.flatnonzero(np.any(np.stack([tree[k][:n] != ref[k][:n] for k in KEY], axis=1), axis=1))
        # Renewed leaves depend on the device tree's own partition: checked on every tree.
        worst = max(worst, _check_renewed_leaves(oracle, ods, tree, g, h, 1.0, f"class tree {t}"))
        if tree.size == ref.size and diff.size == 0:
            identical += 1
        else:
            k = int(diff[0]) if diff.size else n
            assert tree["gain"][k] == pytest.approx(ref["gain"][k], rel=1e-9), f"class tree {t} node {k}: not a near-tie"
            near += 1
        margins[:, c] = margins[:, c] + lr * oracle.tree_predict(ods, tree, ods.features)
    print(f"softmax 4M x 64, K = 10: {identical} of {K} class trees identical, {near} near-ties; renewed leaves "
          f"within {worst:.2e} of the exact solve", flush=True)
    assert identical + near == K and identical >= K - 2


def _stochastic_round(g, seed, qmax):
    """The device's rounding of unit-Hessian gradients: multiplier qmax / max|g|, floor(g m + u)
    with u from splitmix64 keyed by (round seed, 2 * row)."""

    return _round_channel(g, seed, 0, qmax)


@pytest.mark.parametrize("form,qmax", [("int16", 32767.0), ("int8", 127.0)])
def test_config2_renewed_leaves_on_prefix(backend, oracle, config1, form, qmax):
    """The quantised forms with leaf renewal on the first 1,000,000 rows of the 10M x 128 data:
    tree by tree on the device's predictions, the structure equals the oracle's grown from the
    same stochastically rounded gradients (replayed) up to near-ties, and every leaf equals the
    exact solve from the unrounded FP64 gradients of its rows."""

    X, y, _ = config1
    n, trees, seed = 1_000_000, 4, 7
    Xp, yp = np.ascontiguousarray(X[:n]), y[:n]
    ods = oracle.quantize(Xp, 256)
    ds = N._quantize_dense_device(backend.session, Xp, 256)
    model = _fit(backend, ds, yp, form, trees=trees, near_ties="keep")  # the rounded data's own choices
    st = model.fit_stats()
    assert st["renewed_trees"] == trees
    p = np.full(n, model.base_score)
    identical = near = 0
    worst = 0.0
    for t, tree in enumerate(model.trees):
        g = p - yp
        s_t = (seed + t * 0x9E3779B97F4A7C15) % (1 << 64)
        pairs = np.stack([_stochastic_round(g, s_t, qmax), np.ones(n)], axis=1)
        ref = oracle.train_tree(ods, pairs, max_depth=C1["depth"], min_samples_leaf=C1["msl"],
                                min_child_weight=C1["mcw"], reg_lambda=1.0)
        worst = max(worst, _check_renewed_leaves(oracle, ods, tree, g, np.ones(n), 1.0, f"{form} tree {t}"))
        m = min(tree.size, ref.size)
        diff = np.flatnonzero(np.any(np.stack([tree[k][:m] != ref[k][:m] for k in KEY], axis=1), axis=1))
        if tree.size == ref.size and diff.size == 0:
            identical += 1
        else:
            k = int(diff[0]) if diff.size else m
            assert tree["gain"][k] == pytest.approx(ref["gain"][k], rel=1e-9), f"{form} tree {t} node {k}: not a near-tie"
            near += 1
        p = p + C1["lr"] * oracle.tree_predict(ods, tree, ods.features)
    print(f"1M-row prefix {form}: {identical} of {trees} trees identical to the replayed-rounding oracle, {near} "
          f"near-ties; renewed leaves within {worst:.2e} of the exact solve", flush=True)
    assert identical + near == trees
    assert np.array_equal(model.predict_binned(backend.session, ds), p)
