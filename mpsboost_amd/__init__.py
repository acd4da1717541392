"""MI355X (gfx950) histogram-boosting training backend.

The device library is ``mpsboost_amd/lib/libmpsboost_hip.so`` (C ABI in
``include/mpsboost_hip.h``); ``mpsboost_amd._native`` binds it with the reference's
``_native`` names (``_train_regressor_hip``, ``_train_multiclass_softmax_hip``,
``_HipBackend`` test hooks).
"""

from . import _native
from ._native import BackendError, DataError, TrainingError

__all__ = ["BackendError", "DataError", "TrainingError", "is_available", "_native"]
__version__ = "0.1.0"


def is_available() -> bool:
    """True when the library is built and a gfx950 device is visible."""

    try:
        return bool(_native._backend_info()["available"])
    except BackendError:
        return False
