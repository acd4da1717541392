"""Device binning (QuantizeDense / TransformDense of src/core/binned_dataset.cpp:50-238 on the
GPU) against the host binning, which is itself pinned to the oracle and the compiled reference.

Bar: integer work, so bit-exact: bin counts, boundaries (as float32 bit patterns), missing
counts, the feature-major bins and missing mask, and the row-major device codes. The one
documented exception: a zero boundary of a column holding both -0.0 and +0.0 may differ in
sign (compared by value there; the bins must still be identical).
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = pytest.importorskip("mpsboost_amd._native")


@pytest.fixture(scope="module")
def session():
    return N._HipBackend().session


def _same(a, b, zero_sign=False):
    assert (a.rows, a.features, a.max_bins) == (b.rows, b.features, b.max_bins)
    assert np.array_equal(a.bin_counts, b.bin_counts)
    assert np.array_equal(a.missing_counts, b.missing_counts)
    if zero_sign:
        assert np.array_equal(a.boundary_values, b.boundary_values)  # by value: -0.0 == +0.0
    else:
        assert np.array_equal(a.boundary_values.view(np.uint32), b.boundary_values.view(np.uint32))
    assert (a.code_bytes, a.row_stride, a.cells) == (b.code_bytes, b.row_stride, b.cells)
    assert np.array_equal(a.bins, b.bins)
    assert np.array_equal(a.missing, b.missing)


def _matrix(seed, rows, feats, dtype=np.float32):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats))
    if feats > 2:
        X[:, 1] = rng.integers(0, 7, rows)          # few distinct values
    if feats > 3:
        X[:, 3] = 2.5                                # constant
    if feats > 4:
        X[:, 4] = np.nan                             # all missing
    if feats > 5:
        X[:, 5] = rng.lognormal(0.0, 3.0, rows)      # wide dynamic range
    X[rng.random(X.shape) < 0.05] = np.nan
    return X.astype(dtype)


SHAPES = [(1, 1), (2, 3), (63, 5), (65, 64), (1000, 70), (4099, 130), (100_003, 9)]


@pytest.mark.parametrize("rows,feats", SHAPES)
@pytest.mark.parametrize("max_bins", [2, 16, 256])
def test_device_quantize_matches_host(session, rows, feats, max_bins):
    X = _matrix(rows * 7 + feats, rows, feats)
    _same(N._quantize_dense_device(session, X, max_bins), N._quantize_dense(X, max_bins))


@pytest.mark.parametrize("max_bins", [257, 1000, 4096, 65536])
def test_device_quantize_wide_bins(session, max_bins):
    """Two-byte codes; boundaries searched in LDS up to 48 KB per feature group, beyond that
    in global memory."""

    rng = np.random.default_rng(max_bins)
    X = rng.standard_normal((70_000, 20)).astype(np.float32)
    X[:, 7] = np.round(X[:, 7] * 100) + 0.0  # many ties; + 0.0 turns -0.0 into +0.0
    X[rng.random(X.shape) < 0.02] = np.nan
    _same(N._quantize_dense_device(session, X, max_bins), N._quantize_dense(X, max_bins))


def test_device_quantize_float64_and_views(session):
    X = _matrix(5, 3000, 12, np.float64)
    X[7, 2] = 3.4028234663852886e38  # largest float32 exactly: in range
    for view in (X, np.asfortranarray(X), X[::3, ::2], X[:, 1:]):
        _same(N._quantize_dense_device(session, view, 64), N._quantize_dense(view, 64))
    raw = np.zeros(3000 * 12 * 8 + 4, dtype=np.uint8)  # float64 rows at an odd byte offset
    odd = np.ndarray((3000, 12), dtype=np.float64, buffer=raw, offset=4, strides=(96, 8))
    odd[:] = X
    _same(N._quantize_dense_device(session, odd, 64), N._quantize_dense(odd, 64))
    # float32 rows 50 bytes apart: element addresses not 4-byte aligned (byte-wise reads).
    Y = X.astype(np.float32)
    raw = np.zeros(3000 * 50, dtype=np.uint8)
    packed = np.ndarray((3000, 12), dtype=np.float32, buffer=raw, offset=0, strides=(50, 4))
    packed[:] = Y
    _same(N._quantize_dense_device(session, packed, 64), N._quantize_dense(packed, 64))


def test_device_quantize_matches_oracle(session):
    X = _matrix(9, 5000, 8)
    a, b = N._quantize_dense_device(session, X, 32), O.oracle().quantize(X, 32)
    assert np.array_equal(a.bin_counts, b.bin_counts)
    assert np.array_equal(a.bins, b.bins) and np.array_equal(a.missing, b.missing)
    assert a.boundaries == b.boundaries


def test_signed_zeros(session):
    rng = np.random.default_rng(3)
    X = rng.choice(np.array([-0.0, 0.0, 1.0, -1.0, 2.0], dtype=np.float32), size=(4000, 3))
    _same(N._quantize_dense_device(session, X, 4), N._quantize_dense(X, 4), zero_sign=True)


def test_device_transform_matches_host(session):
    X = _matrix(11, 6000, 10)
    Z = _matrix(12, 2500, 10) * 2.0
    Z[::7, 0] = np.nan
    X[:, 0] = np.nan_to_num(X[:, 0])  # feature 0: no missing at training time, NaN in Z
    schema = N._quantize_dense(X, 64)
    _same(N._transform_dense_device(session, Z, schema), N._transform_dense(Z, schema))
    _same(N._transform_dense_device(session, Z.astype(np.float64), schema), N._transform_dense(Z, schema))


@pytest.mark.parametrize("first", ["inf", "range"])
def test_device_binning_errors(session, first):
    X = np.zeros((6, 4))
    inf_at, big_at = ((4, 1), (0, 2)) if first == "inf" else ((0, 3), (5, 1))
    X[inf_at] = np.inf
    X[big_at] = 1e300
    with pytest.raises(N.DataError) as host:
        N._quantize_dense(X, 8)
    with pytest.raises(N.DataError) as dev:
        N._quantize_dense_device(session, X, 8)
    assert str(dev.value) == str(host.value)
    with pytest.raises(N.DataError, match="Inf"):
        N._quantize_dense_device(session, np.array([[1.0, -np.inf]], dtype=np.float32), 8)
    with pytest.raises(N.DataError, match="max_bins"):
        N._quantize_dense_device(session, np.zeros((4, 2), dtype=np.float32), 1)
    schema = N._quantize_dense(np.zeros((6, 3), dtype=np.float32), 8)
    with pytest.raises(N.DataError, match="feature count"):
        N._transform_dense_device(session, np.zeros((6, 4), dtype=np.float32), schema)
    # The session still works after the errors.
    Y = _matrix(1, 500, 6)
    _same(N._quantize_dense_device(session, Y, 16), N._quantize_dense(Y, 16))


def test_fit_on_device_binned_data_is_identical(session):
    """The device-binned dataset is resident; a fit on it equals the fit on host-binned data and
    the oracle's (row-order FP64 form: bit-identical)."""

    X = _matrix(21, 8000, 7)
    rng = np.random.default_rng(4)
    y = np.nan_to_num(X[:, 0]) * 2 + np.nan_to_num(X[:, 5]) * 0.1 + rng.standard_normal(8000) * 0.1
    p = N._TrainingParameters(4, 0.3, 64, 5, 5, 0.0, 1.0, precision="fp64")
    dev = N._quantize_dense_device(session, X, 64)
    m_dev = N._train_binned(session, dev, y, None, p)
    host = N._quantize_dense(X, 64)
    m_host = N._train_binned(session, host, y, None, p)
    assert m_dev.serialize() == m_host.serialize()
    # Once the session's code buffer holds other data, and in another session, a fit on the
    # device-binned dataset uploads its codes again.
    assert N._train_binned(session, dev, y, None, p).serialize() == m_dev.serialize()
    assert N._train_binned(N._Session(0), dev, y, None, p).serialize() == m_dev.serialize()
    assert np.array_equal(dev.bins, host.bins) and np.array_equal(dev.missing, host.missing)
    o = O.oracle()
    ref = o.train_regressor(o.quantize(X, 64), y, n_estimators=4, learning_rate=0.3, max_depth=5,
                            min_samples_leaf=5)
    for a, b in zip(m_dev.trees, ref.trees):
        for k in a.dtype.names:
            assert np.array_equal(a[k], b[k]), k
    Z = _matrix(22, 3000, 7)
    tz = N._transform_dense_device(session, Z, dev)
    assert np.array_equal(m_dev.predict_binned(session, tz), ref.predict(o.transform(Z, o.quantize(X, 64))))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sampled_binning_from_resident_matrix(session, dtype):
    """matrix[rows][:, features] binned in place equals binning the materialised sub-matrix
    (rows repeat as in bootstrap samples; features in sampled order)."""

    X = _matrix(31, 20_000, 40, dtype)
    dm = N._DeviceMatrix(X)
    rng = np.random.default_rng(2)
    for n_rows, n_feats, mb in ((20_000, 40, 256), (7_001, 9, 16), (50, 3, 300)):
        rows = rng.integers(0, 20_000, n_rows)
        feats = rng.choice(40, n_feats, replace=False)
        _same(N._quantize_matrix_sampled(session, dm, rows, feats, mb), N._quantize_dense(X[rows][:, feats], mb))
    view = X[:, ::3]  # a strided view stays strided on the device
    dv = N._DeviceMatrix(view)
    rows = rng.integers(0, 20_000, 3000)
    feats = np.array([5, 0, 13])
    _same(N._quantize_matrix_sampled(session, dv, rows, feats, 32), N._quantize_dense(view[rows][:, feats], 32))


def test_sampled_binning_errors(session):
    X = _matrix(32, 500, 6)
    X[7, 2] = np.inf
    dm = N._DeviceMatrix(X)
    rows = np.arange(100, 300)
    _same(N._quantize_matrix_sampled(session, dm, rows, np.array([0, 2, 4]), 16),
          N._quantize_dense(X[rows][:, [0, 2, 4]], 16))  # the Inf is outside the sample
    with pytest.raises(N.DataError, match="Inf"):
        N._quantize_matrix_sampled(session, dm, np.array([3, 7, 9]), np.array([1, 2]), 16)
    with pytest.raises(ValueError, match="out of range"):
        N._quantize_matrix_sampled(session, dm, np.array([500]), np.array([1]), 16)
    with pytest.raises(ValueError, match="out of range"):
        N._quantize_matrix_sampled(session, dm, np.array([1]), np.array([6]), 16)
