"""Model files in the reference's MPSBMOD 1.4 format (src/io/model_format.cpp,
src/io/model_payload.cpp, src/io/model_file_io.cpp).

Host-only checks compare this library's reader/writer with the reference's own, compiled from
its sources (oracle/_ref/libref.so): bytes written by the reference round-trip unchanged, and
every corruption is rejected with the same error class and message. The GPU check trains in the
row-order FP64 form and requires the serialized model to equal, byte for byte, the reference's
serialization of the reference-trained model."""

import struct

import numpy as np
import pytest

from oracle import oracle as O

N = pytest.importorskip("mpsboost_amd._native")
REF = O.reference()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref/libref.so was not built")

FIELDS = ["feature_index", "threshold_bin", "left_child", "right_child", "default_left", "flags", "leaf_value",
          "gain"]
HEADER = 32


def _data(seed=3, rows=2000, feats=5):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, feats)).astype(np.float32)
    X[rng.random(X.shape) < 0.05] = np.nan
    y = np.sin(np.nan_to_num(X[:, 0])) + np.nan_to_num(X[:, 1]) * np.nan_to_num(X[:, 2])
    return X, y


def _fnv1a(data: bytes) -> int:
    h = 14695981039346656037
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _reframe(blob: bytes, payload: bytes, minor=None) -> bytes:
    """New header (size and checksum) around a modified payload."""

    major, old_minor, kind = struct.unpack_from("<HHI", blob, 8)
    return blob[:8] + struct.pack("<HHIQQ", major, old_minor if minor is None else minor, kind, len(payload),
                                  _fnv1a(payload)) + payload


def _reference_regression_bytes():
    X, y = _data()
    ds = REF.quantize(X, 32)
    return REF.train_regressor(ds, y, n_estimators=3, learning_rate=0.3, max_depth=3, min_samples_leaf=5).serialize()


def _reference_multiclass_bytes():
    X, _ = _data(5)
    y = np.random.default_rng(1).integers(0, 3, X.shape[0]).astype(np.float64)
    ds = REF.quantize(X, 16)
    return REF.train_multiclass(ds, y, 3, n_estimators=2, learning_rate=0.3, max_depth=2, min_samples_leaf=5).serialize()


@needs_ref
@pytest.mark.parametrize("kind", ["regression", "multiclass"])
def test_reference_bytes_round_trip(kind):
    blob = _reference_regression_bytes() if kind == "regression" else _reference_multiclass_bytes()
    ours = N._deserialize_model(blob)
    assert ours.serialize() == blob
    theirs = REF.deserialize_model(blob)
    assert ours.tree_count == theirs.tree_count
    for a, b in zip(ours.trees, theirs.trees):
        for k in FIELDS:
            assert np.array_equal(a[k], b[k])


def _error_of(load, blob):
    try:
        load(blob)
    except (N.DataError, N.TrainingError, O.DataError, O.TrainingError) as e:
        return ("data" if isinstance(e, (N.DataError, O.DataError)) else "training", str(e))
    return None


def _corruptions(blob):
    payload = blob[HEADER:]
    node0 = _first_node_offset(blob)  # tree 0, node 0 within the payload
    out = {
        "truncated header": blob[:20],
        "bad magic": b"X" + blob[1:],
        "major version": blob[:8] + struct.pack("<H", 2) + blob[10:],
        "payload length": blob + b"\x00",
        "checksum": blob[:-1] + bytes([blob[-1] ^ 1]),
        "trailing bytes": _reframe(blob, payload + b"\x00"),
        "truncated payload": _reframe(blob, payload[:-5]),
    }
    p = bytearray(payload)
    p[node0 + 34] = 1  # first reserved byte of node 0
    out["reserved byte"] = _reframe(blob, bytes(p))
    p = bytearray(payload)
    struct.pack_into("<I", p, node0 + 8, 0)  # node 0's left child points back at itself
    out["cycle"] = _reframe(blob, bytes(p))
    p = bytearray(payload)
    struct.pack_into("<I", p, node0 + 4, 10_000)  # threshold beyond the feature's bins
    out["threshold"] = _reframe(blob, bytes(p))
    p = bytearray(payload)
    struct.pack_into("<d", p, node0 + 24, float("nan"))  # gain of node 0
    out["non-finite gain"] = _reframe(blob, bytes(p))
    p = bytearray(payload)
    struct.pack_into("<f", p, _first_boundary_offset(blob) + 4, -1e30)  # second boundary below the first
    out["unsorted boundaries"] = _reframe(blob, bytes(p))
    return out


def _first_boundary_offset(blob):
    features = struct.unpack_from("<I", blob, HEADER)[0]
    return 4 + 4 + 8 + 8 + 4 + 8 + 8 + 8 + features * 24


def _first_node_offset(blob):
    payload = blob[HEADER:]
    total = struct.unpack_from("<Q", payload, 44)[0]
    return _first_boundary_offset(blob) + 4 * total + 4 + 4


@needs_ref
def test_corrupt_files_fail_like_the_reference():
    blob = _reference_regression_bytes()
    for name, bad in _corruptions(blob).items():
        ours = _error_of(N._deserialize_model, bad)
        theirs = _error_of(REF.deserialize_model, bad)
        assert theirs is not None, f"{name}: the reference accepted it"
        assert ours == theirs, f"{name}: {ours} vs reference {theirs}"


@needs_ref
def test_older_minor_without_quantile_fields():
    """Minor 3 has no quantile alpha / Tweedie power fields (src/io/model_payload.cpp:214-220)."""

    blob = _reference_regression_bytes()
    payload = blob[HEADER:]
    old = payload[:28] + payload[44:]  # drop objective_alpha and tweedie_variance_power
    v3 = _reframe(blob, old, minor=3)
    ours, theirs = N._deserialize_model(v3), REF.deserialize_model(v3)
    for a, b in zip(ours.trees, theirs.trees):
        for k in FIELDS:
            assert np.array_equal(a[k], b[k])
    assert ours.serialize() == blob  # written back at minor 4 with the defaults 0.5 / 1.5


@needs_ref
def test_file_save_load_round_trip(tmp_path):
    blob = _reference_regression_bytes()
    m = N._deserialize_model(blob)
    path = tmp_path / "model.mpsb"
    m.save(str(path))
    assert path.read_bytes() == blob
    assert N._load_model(str(path)).serialize() == blob
    assert [p.name for p in tmp_path.iterdir()] == ["model.mpsb"]  # no temporary left behind
    with pytest.raises(N.TrainingError, match="directory does not exist"):
        m.save(str(tmp_path / "missing" / "model.mpsb"))
    with pytest.raises(N.TrainingError):
        N._load_model(str(tmp_path / "absent.mpsb"))


@needs_ref
@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["regression", "multiclass"])
def test_device_model_file_equals_reference_file(kind):
    """The row-order FP64 form on the device writes the reference's bytes for the same fit
    (multiclass: structure only, see below)."""

    session = N._HipBackend().session
    if kind == "regression":
        X, y = _data(11, 3000, 6)
        ds, rds = N._quantize_dense(X, 32), REF.quantize(X, 32)
        p = N._TrainingParameters(4, 0.3, 32, 4, 5, 0.0, 1.0, precision="fp64")
        blob = N._train_binned(session, ds, y, None, p).serialize()
        ref = REF.train_regressor(rds, y, n_estimators=4, learning_rate=0.3, max_depth=4, min_samples_leaf=5)
        assert blob == ref.serialize()
        return
    # Softmax leaves use the device's exp, so bytes may differ in the leaves' last bits; the
    # file must still load in the reference and carry the same tree structure.
    X, _ = _data(13, 3000, 6)
    y = np.random.default_rng(2).integers(0, 3, X.shape[0]).astype(np.float64)
    ds, rds = N._quantize_dense(X, 32), REF.quantize(X, 32)
    p = N._TrainingParameters(2, 0.3, 32, 3, 5, 0.0, 1.0, precision="fp64")
    blob = N._train_binned(session, ds, y, None, p, class_count=3).serialize()
    theirs = REF.deserialize_model(blob)
    ref = REF.train_multiclass(rds, y, 3, n_estimators=2, learning_rate=0.3, max_depth=3, min_samples_leaf=5)
    for a, b in zip(theirs.trees, ref.trees):
        for k in FIELDS[:6]:
            assert np.array_equal(a[k], b[k])
        np.testing.assert_allclose(a["leaf_value"], b["leaf_value"], rtol=1e-12, atol=1e-15)
