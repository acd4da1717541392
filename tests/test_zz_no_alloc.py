"""Allocation-free histogram build (mpsb_build_histograms_noalloc) against the entry point it is
a variant of, mpsb_build_histograms, called through the C ABI on the same inputs.

The test allocates the device inputs and outputs itself, through the HIP runtime the library
is linked against (loaded by its soname after the library, so both use one runtime); the
allocation-free call uses only its fixed scratch array. The calls run on a stream the test
creates. Two calls in a row, the second on a smaller dataset and node, so the
second runs on scratch that still holds the first call's contents. Bars: the int16, int8 and
fp64_fixed forms bit-identical (same rounding stream, exact integer sums); fp32 counts exact
and, since the in-block FP32 addition order is not fixed, each entry point's sums within the
worst-case bound of FP32 recursive summation of the cell's values (Higham, Accuracy and
Stability of Numerical Algorithms, 2nd ed., section 4.2: |error| <= gamma_(n+1) * sum |x|, with
n the cell's count and one more rounding for the FP64 -> FP32 conversion of each value) of the
exact FP64 sums.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libmpsboost_hip.so already loaded."""

    def __init__(self):
        N._lib()  # the library first: its libamdhip64.so.7 is the one dlopen returns below
        self.rt = C.CDLL("libamdhip64.so.7")
        self.stream = C.c_void_p()
        self._ok(self.rt.hipStreamCreate(C.byref(self.stream)))
        self.live = []

    def _ok(self, rc):
        assert rc == 0, f"HIP runtime call failed with {rc}"

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 1))))
        self.live.append(p)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1))
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        self._ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2))
        return out

    def free(self):
        self._ok(self.rt.hipDeviceSynchronize())
        for p in self.live:
            self._ok(self.rt.hipFree(p))
        self.live = []


@pytest.fixture(scope="module")
def hip():
    h = _Hip()
    yield h
    h.free()
    h.rt.hipStreamDestroy(h.stream)


def _case(seed, rows, feats, unit_h, node_every):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    X[:, 1] = rng.integers(0, 10, rows)  # a narrow feature
    X[rng.random(X.shape) < 0.05] = np.nan
    ds = N._quantize_dense(X, 256)
    pairs = np.empty((rows, 2))
    pairs[:, 0] = rng.standard_normal(rows)
    pairs[:, 1] = 1.0 if unit_h else rng.uniform(0.1, 2.0, rows)
    node = np.arange(seed % node_every, rows, node_every, dtype=np.uint64)
    return ds, pairs, node


def _noalloc(hip, ds, pairs, node, precision):
    n_ref = int(ds.bin_counts.sum())
    sentinel_c = np.full(n_ref, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    sentinel_f = np.full(n_ref, np.nan)
    codes = hip.upload(ds.device_codes())
    rows = hip.upload(node.astype(np.uint64))
    gh = hip.upload(pairs.reshape(-1))
    cnt, g, h = hip.upload(sentinel_c), hip.upload(sentinel_f), hip.upload(sentinel_f)
    try:
        N._build_histograms_noalloc(ds, codes, rows, node.size, gh, precision, cnt, g, h, hip.stream.value)
        return hip.download(cnt, sentinel_c), hip.download(g, sentinel_f), hip.download(h, sentinel_f)
    finally:
        hip.free()


def _exact(ds, pairs, node):
    """FP64 per-cell sums of g, |g|, h, |h| in the reference view (bin 0 holds the missing rows)."""

    out = [[], [], [], []]
    for f in range(ds.features):
        b = ds.bins[f, node.astype(np.int64)]
        nb = int(ds.bin_counts[f])
        for k, v in enumerate((pairs[node.astype(np.int64), 0], pairs[node.astype(np.int64), 1])):
            out[2 * k].append(np.bincount(b, weights=v, minlength=nb))
            out[2 * k + 1].append(np.bincount(b, weights=np.abs(v), minlength=nb))
    return [np.concatenate(x) for x in out]


@pytest.mark.parametrize("precision", ["int16", "int8", "fp64_fixed", "fp32"])
@pytest.mark.parametrize("unit_h", [True, False])
def test_allocation_free_matches_mpsb_build_histograms(hip, precision, unit_h):
    backend = N._HipBackend()
    for seed, rows, feats, every in ((31, 200_000, 20, 2), (32, 30_000, 9, 6)):  # second call smaller
        print(f"noalloc: {precision} unit_h={unit_h} {rows}x{feats} node every {every}", flush=True)
        ds, pairs, node = _case(seed, rows, feats, unit_h, every)
        c, g, h = _noalloc(hip, ds, pairs, node, precision)
        rc, rg, rh = backend.histogram_arrays(ds, pairs, node, precision)
        assert np.array_equal(c, rc)
        assert int(c.sum()) == node.size * feats
        if precision == "fp32":
            eg, ag, eh, ah = _exact(ds, pairs, node)
            gamma = (c + 1.0) * 2.0**-24 / (1.0 - (c + 1.0) * 2.0**-24)
            for got in (g, rg):
                assert np.all(np.abs(got - eg) <= gamma * ag)
            for got in (h, rh):
                assert np.all(np.abs(got - eh) <= gamma * ah)
        else:
            assert np.array_equal(g, rg) and np.array_equal(h, rh)


def test_allocation_free_errors_match(hip):
    """A row id past the dataset and a non-finite gradient give the existing entry point's errors."""

    ds, pairs, node = _case(33, 5_000, 4, True, 3)
    bad_rows = node.copy()
    bad_rows[-1] = 5_000
    with pytest.raises(N.TrainingError, match="Histogram row index is out of bounds"):
        _noalloc(hip, ds, pairs, bad_rows, "int16")
    bad_pairs = pairs.copy()
    bad_pairs[7, 0] = np.inf
    with pytest.raises(N.TrainingError, match="finite") as got:
        _noalloc(hip, ds, bad_pairs, node, "int16")
    with pytest.raises(N.TrainingError) as ref:
        N._HipBackend().histogram_arrays(ds, bad_pairs, node, "int16")
    assert str(got.value) == str(ref.value)
    c, g, h = _noalloc(hip, ds, pairs, node, "int16")  # a good call after the failures
    rc, rg, rh = N._HipBackend().histogram_arrays(ds, pairs, node, "int16")
    assert np.array_equal(c, rc) and np.array_equal(g, rg) and np.array_equal(h, rh)
