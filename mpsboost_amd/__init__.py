"""MI355X (gfx950) histogram-boosting training backend.

The device library is ``mpsboost_amd/lib/libmpsboost_hip.so`` (C ABI in
``include/mpsboost_hip.h``); ``mpsboost_amd._native`` binds it with the reference's
``_native`` names (``_train_regressor_hip``, ``_train_multiclass_softmax_hip``,
``_load_regression_model``, ``_HipBackend`` test hooks). ``mpsboost_amd.estimators`` holds the
scikit-learn-style estimators with ``device="hip"`` and ``mpsboost_amd.device_policy`` the
device choice.
"""

from . import _native
from ._native import BackendError, DataError, TrainingError
from .device_policy import DeviceDecision, choose_device, is_available
from .estimators import (
    DecisionTreeClassifier,
    DecisionTreeRegressor,
    ExtraTreeClassifier,
    ExtraTreeRegressor,
    ExtraTreesClassifier,
    ExtraTreesRegressor,
    GradientBoostingClassifier,
    GradientBoostingRegressor,
    MPSBoostClassifier,
    MPSBoostRegressor,
    NotFittedError,
    RandomForestClassifier,
    RandomForestRegressor,
)

__all__ = [
    "BackendError", "DataError", "DecisionTreeClassifier", "DecisionTreeRegressor", "DeviceDecision",
    "ExtraTreeClassifier", "ExtraTreeRegressor", "ExtraTreesClassifier", "ExtraTreesRegressor",
    "GradientBoostingClassifier", "GradientBoostingRegressor", "MPSBoostClassifier", "MPSBoostRegressor",
    "NotFittedError", "RandomForestClassifier", "RandomForestRegressor", "TrainingError", "choose_device",
    "is_available", "_native",
]
__version__ = "0.1.0"
