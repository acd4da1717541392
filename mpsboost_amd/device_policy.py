"""Which device a fit runs on, and why.

Behaviour (the interface of the reference's src/mpsboost/device_policy.py:52-99, with the HIP
device in the place of Metal): an explicit "cpu" or "hip" request is honoured as it is; "auto"
takes the HIP device when one is present and the estimated histogram work (rows x features x
trees x bins) reaches the threshold, and otherwise the CPU trainer. This package ships no CPU
trainer, so by default (``cpu_available=False``) a small "auto" fit still goes to the HIP device,
and without a device it fails loudly; a host package that has a CPU trainer passes
``cpu_available=True`` and gets the threshold rule. The reason strings, the default threshold,
the argument checks with their messages and the keys of ``decision_to_dict`` are the
reference's, so diagnostics read the same.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, fields
from typing import Literal

RequestedDevice = Literal["cpu", "hip", "auto"]
ResolvedDevice = Literal["cpu", "hip"]

DEFAULT_WORK_THRESHOLD = 134_217_728  # 2**27 histogram updates

# Argument checks in the order they are applied (the first failing one raises).
_WORK_TERMS = ("n_samples", "n_features", "n_estimators", "max_bins")


@dataclass(frozen=True)
class DeviceDecision:
    """One backend-selection decision and its reason."""

    requested: RequestedDevice
    selected: ResolvedDevice
    reason: str
    estimated_work: int
    hip_available: bool


def _strict_int(name: str, value: object) -> int:
    # bool is an int subclass, and True must not count as a row count.
    if type(value) is bool or not isinstance(value, int):
        raise TypeError(f"{name} must be an integer")
    return value


def _estimated_histogram_work(**terms: int) -> int:
    """Histogram updates of a fit: the product of the four terms, each a positive integer.
    Python integers do not overflow, so the product is exact at any size."""

    checked = []
    for name in _WORK_TERMS:
        value = _strict_int(name, terms[name])
        if value < 1:
            raise ValueError(f"{name} must be positive")
        checked.append(value)
    return math.prod(checked)


@dataclass(frozen=True)
class _AutoRule:
    threshold: int
    cpu_trainer: bool

    def resolve(self, work: int, available: bool) -> tuple[str, str]:
        if not available:
            return "cpu", "hip unavailable"
        if work >= self.threshold:
            return "hip", "estimated work meets hip threshold"
        if self.cpu_trainer:
            return "cpu", "estimated work below hip threshold"
        return "hip", "estimated work below hip threshold; no cpu trainer in this package"


def choose_device(*, requested: RequestedDevice, n_samples: int, n_features: int, n_estimators: int, max_bins: int,
                  hip_available: bool, hip_work_threshold: int = DEFAULT_WORK_THRESHOLD,
                  cpu_available: bool = False) -> DeviceDecision:
    """Choose the backend for one fit and explain the decision."""

    if requested not in ("cpu", "hip", "auto"):
        raise ValueError("requested device must be 'cpu', 'hip', or 'auto'")
    if type(hip_available) is not bool:
        raise TypeError("hip_available must be a bool")
    threshold = _strict_int("hip_work_threshold", hip_work_threshold)
    if threshold < 1:
        raise ValueError("hip_work_threshold must be positive")
    work = _estimated_histogram_work(n_samples=n_samples, n_features=n_features, n_estimators=n_estimators,
                                     max_bins=max_bins)
    if requested == "auto":
        selected, reason = _AutoRule(threshold, bool(cpu_available)).resolve(work, hip_available)
    else:
        selected, reason = requested, f"explicit {requested} request"
    return DeviceDecision(requested, selected, reason, work, hip_available)


def decision_to_dict(decision: DeviceDecision) -> dict[str, object]:
    """A stable, serialisable record of a decision (plain str / int / bool values)."""

    casts = {"estimated_work": int, "hip_available": bool}
    return {f.name: casts.get(f.name, str)(getattr(decision, f.name)) for f in fields(decision)}


def is_available() -> bool:
    """Whether a supported (gfx950) HIP device is present; never raises for its absence."""

    from . import _native

    try:
        return bool(_native._backend_info()["available"])
    except _native.BackendError:
        return False


def setup_instructions() -> str:
    """What to do when the HIP device is missing."""

    return ("No supported HIP device was found: this package needs an AMD Instinct MI355X (gfx950) with ROCm, "
            "and libmpsboost_hip.so built by `python -m mpsboost_amd.build`.")
