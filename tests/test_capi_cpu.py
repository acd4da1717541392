"""The C ABI on the host: the library loads and exports every entry point the header declares;
host-side work (binning, schema transform, argument validation) matches the oracle and the
reference's messages; device entry points fail cleanly without a device. No device compute."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

N = pytest.importorskip("mpsboost_amd._native")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpsboost_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpsb_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N._lib()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in mpsboost_hip.h but not exported: {missing}"


def test_exported_symbols_are_exactly_the_c_abi():
    """nm -D: the library's own exported functions are the header's, nothing C++ leaks."""

    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", N.LIBRARY], capture_output=True, text=True, check=True).stdout
    mine = sorted({line.split()[-1] for line in out.splitlines() if line.split()[-1].startswith("mpsb_")})
    assert mine == _declared()


def test_backend_info_is_a_value_not_an_error():
    info = N._backend_info()
    assert isinstance(info["available"], bool)
    if not info["available"]:
        assert info["device_name"] is None


def _device_present():
    return N._backend_info()["available"]


def test_session_without_device_is_backend_error():
    if _device_present():
        pytest.skip("a device is visible")
    with pytest.raises(N.BackendError, match="gfx950|device"):
        N._Session(0)


@pytest.mark.parametrize("max_bins", [2, 16, 256, 300])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_quantize_matches_oracle(max_bins, dtype):
    rng = np.random.default_rng(max_bins)
    X = rng.standard_normal((3001, 7)).astype(dtype)
    X[:, 2] = rng.integers(0, 5, 3001)
    X[:, 4] = 1.5  # constant column
    X[rng.random(X.shape) < 0.07] = np.nan
    X[:, 6] = np.nan  # all missing
    o = O.oracle()
    a, b = N._quantize_dense(X, max_bins, threads=3), o.quantize(X, max_bins)
    assert np.array_equal(a.bin_counts, b.bin_counts)
    assert np.array_equal(a.bins, b.bins) and np.array_equal(a.missing, b.missing)
    assert a.boundaries == b.boundaries
    assert a.bin_width == (8 if max_bins <= 256 else 16)


def test_quantize_strided_view_matches_contiguous():
    rng = np.random.default_rng(4)
    base = rng.standard_normal((500, 12)).astype(np.float32)
    view = base[::2, ::3]
    a, b = N._quantize_dense(view, 32), N._quantize_dense(np.ascontiguousarray(view), 32)
    assert np.array_equal(a.bins, b.bins) and a.boundaries == b.boundaries
    Xf = np.asfortranarray(base)
    assert np.array_equal(N._quantize_dense(Xf, 32).bins, N._quantize_dense(base, 32).bins)


def test_transform_uses_frozen_schema():
    rng = np.random.default_rng(9)
    X = rng.standard_normal((2000, 5)).astype(np.float32)
    Z = rng.standard_normal((700, 5)).astype(np.float32) * 3.0
    Z[::5, 1] = np.nan
    o = O.oracle()
    sa, sb = N._quantize_dense(X, 64), o.quantize(X, 64)
    ta, tb = N._transform_dense(Z, sa), o.transform(Z, sb)
    assert np.array_equal(ta.bins, tb.bins) and np.array_equal(ta.missing, tb.missing)
    assert ta.boundaries == sa.boundaries


def test_dataset_from_bins_round_trip():
    rng = np.random.default_rng(2)
    bins = rng.integers(0, 9, (4, 1000)).astype(np.uint8)
    missing = np.zeros_like(bins)
    missing[1, ::7] = 1
    bins[1, ::7] = 0
    ds = N._dataset_from_bins(bins, [9, 9, 9, 9], 16, missing)
    assert np.array_equal(ds.bins, bins) and np.array_equal(ds.missing, missing.astype(bool))
    assert ds.cells == 4 * 9 + 1  # one missing cell for the feature with missing values


@pytest.mark.parametrize("bad,match", [
    (dict(max_depth=6, min_samples_leaf=0), "min_samples_leaf"),
    (dict(max_depth=6, reg_lambda=-1.0), "reg_lambda"),
    (dict(max_depth=6, reg_alpha=-0.5), "reg_alpha"),
    (dict(max_depth=6, max_delta_step=-1.0), "max_delta_step"),
])
def test_invalid_parameters_fail_before_any_device_work(bad, match):
    """mpsb_train_regressor validates (with the reference's messages) before it touches the
    session, so a NULL session still reports the parameter error."""

    class NoSession:
        handle = None

    X = np.random.default_rng(0).standard_normal((50, 3)).astype(np.float32)
    ds = N._quantize_dense(X, 16)
    y = np.zeros(50)
    kw = dict(bad)
    depth = kw.pop("max_depth")
    msl = kw.pop("min_samples_leaf", 1)
    lam = kw.pop("reg_lambda", 1.0)
    with pytest.raises(N.TrainingError, match=match):
        N._train_binned(NoSession, ds, y, None, N._TrainingParameters(3, 0.1, 16, depth, msl, 0.0, lam, **kw))
    with pytest.raises(ValueError, match="session"):  # valid parameters: only the session is missing
        N._train_binned(NoSession, ds, y, None, N._TrainingParameters(3, 0.1, 16, depth, 1, 0.0, 1.0))


def test_bad_inputs_are_data_errors():
    with pytest.raises(N.DataError, match="max_bins"):
        N._quantize_dense(np.zeros((4, 2), dtype=np.float32), 1)
    with pytest.raises(N.DataError):
        N._quantize_dense(np.array([[1.0, np.inf]], dtype=np.float32), 16)
    with pytest.raises(TypeError):
        N._quantize_dense(np.zeros((4, 2), dtype=np.int32), 16)


def _bad_matrix(first):
    """float64 matrix with an Inf and a value beyond float32 range; `first` names the one met
    first in the reference's feature-major read order (src/core/binned_dataset.cpp:151-157)."""

    X = np.zeros((6, 4))
    inf_at, big_at = ((4, 1), (0, 2)) if first == "inf" else ((0, 3), (5, 1))
    X[inf_at] = np.inf
    X[big_at] = 1e300
    return X


@pytest.mark.parametrize("first", ["inf", "range"])
def test_first_bad_value_in_read_order_is_reported(first):
    X = _bad_matrix(first)
    o = O.oracle()
    with pytest.raises(O.DataError) as theirs:
        o.quantize(X, 8)
    with pytest.raises(N.DataError) as ours:
        N._quantize_dense(X, 8, threads=3)
    assert str(ours.value) == str(theirs.value)
    assert ("Inf" in str(ours.value)) == (first == "inf")
    schema = N._quantize_dense(np.zeros((6, 4)) + np.arange(4), 8)
    with pytest.raises(O.DataError) as theirs:
        o.transform(X, o.quantize(np.zeros((6, 4)) + np.arange(4), 8))
    with pytest.raises(N.DataError) as ours:
        N._transform_dense(X, schema)
    assert str(ours.value) == str(theirs.value)


def test_no_silent_fallback_when_library_missing(monkeypatch):
    monkeypatch.setattr(N, "LIBRARY", os.path.join(ROOT, "does-not-exist.so"))
    monkeypatch.setattr(N, "_lib_handle", None)
    with pytest.raises(N.BackendError, match="not built"):
        N._lib()


def test_adapter_header_matches_reference_interfaces(tmp_path):
    """include/mpsboost_hip_backend.hpp implements the reference's GradientComputer,
    HistogramBuilder and LayerHistogramBuilder (include/mpsboost/backend.hpp): compile it
    against the reference's own headers when they are present."""

    import shutil
    import subprocess

    ref_inc = "/root/reference/include"
    if not os.path.isdir(ref_inc) or shutil.which("g++") is None:
        pytest.skip("reference headers or g++ absent")
    src = tmp_path / "t.cpp"
    src.write_text('#include "mpsboost_hip_backend.hpp"\n'
                   "namespace mpsboost { void use(const HistogramBuilder&, const GradientComputer&,\n"
                   "                              const LayerHistogramBuilder&);\n"
                   "void f() { HipBackend b; use(b, b, b); } }\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", f"-I{ref_inc}",
                    f"-I{os.path.join(ROOT, 'include')}", str(src)], check=True)


def test_device_codes_are_the_row_major_layout():
    """mpsb_dataset_codes: row r, feature f holds the reference bin, or bin_count for missing."""

    rng = np.random.default_rng(5)
    X = rng.standard_normal((300, 7)).astype(np.float32)
    X[rng.random(X.shape) < 0.1] = np.nan
    ds = N._quantize_dense(X, 16)
    codes = ds.device_codes()
    assert codes.shape == (300, ds.row_stride * ds.code_bytes)
    expect = np.where(ds.missing, ds.bin_counts[:, None], ds.bins).T
    assert np.array_equal(codes[:, :7], expect.astype(np.uint8))
    assert not codes[:, 7:].any()  # padding bytes are zero


@pytest.mark.parametrize("precision,rows,match", [
    ("fp64", 10, "fp32, int16, int8 and fp64_fixed"),
    ("int16", 0, "Node row set must not be empty"),
    ("int16", (1 << 20) + 1, "at most 1,048,576"),
])
def test_allocation_free_build_rejects_before_device_work(precision, rows, match):
    """The limits and form checks of mpsb_build_histograms_noalloc run on the host before any
    device call (the addresses are never dereferenced)."""

    X = np.random.default_rng(6).standard_normal((64, 3)).astype(np.float32)
    ds = N._quantize_dense(X, 8)
    with pytest.raises((ValueError, N.TrainingError), match=match):
        N._build_histograms_noalloc(ds, 16, 16, rows, 16, precision, 16, 16, 16)


@pytest.mark.parametrize("field,table", [("near_ties", "NEAR_TIES"), ("leaf_solve", "LEAF_SOLVE"),
                                         ("class_rounds", "CLASS_ROUNDS")])
def test_solver_switches_are_validated(monkeypatch, field, table):
    """The device-only switches (leaf renewal, near-tie resolution) reject unknown names in
    Python and out-of-range codes in mpsb_train_regressor, before any device work."""

    class NoSession:
        handle = None

    X = np.random.default_rng(1).standard_normal((40, 3)).astype(np.float32)
    ds = N._quantize_dense(X, 16)
    y = np.zeros(40)
    with pytest.raises(ValueError, match=field):
        N._TrainingParameters(3, 0.1, 16, 4, 1, 0.0, 1.0, **{field: "bogus"})
    monkeypatch.setitem(getattr(N, table), "bogus", 7)
    with pytest.raises((ValueError, N.TrainingError), match=field):
        N._train_binned(NoSession, ds, y, None, N._TrainingParameters(3, 0.1, 16, 4, 1, 0.0, 1.0, **{field: "bogus"}))
