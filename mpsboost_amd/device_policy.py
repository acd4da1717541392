"""Device selection with the HIP device (the reference's src/mpsboost/device_policy.py:52-99).

Explicit requests are honoured. ``auto`` picks the HIP device when it is available and the
estimated histogram work reaches the threshold; below it, the reference keeps the fit on its
CPU trainer. This package ships no CPU trainer (``cpu_available=False``), so here ``auto`` takes
the HIP device whenever one is present and otherwise fails loudly; a host package that has a
CPU trainer passes ``cpu_available=True`` and gets the reference's threshold behaviour.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

RequestedDevice = Literal["cpu", "hip", "auto"]
ResolvedDevice = Literal["cpu", "hip"]


@dataclass(frozen=True)
class DeviceDecision:
    """One backend-selection decision and its reason."""

    requested: RequestedDevice
    selected: ResolvedDevice
    reason: str
    estimated_work: int
    hip_available: bool


def _estimated_histogram_work(*, n_samples: int, n_features: int, n_estimators: int, max_bins: int) -> int:
    """rows x features x trees x bins, with the reference's argument checks."""

    for name, value in (("n_samples", n_samples), ("n_features", n_features), ("n_estimators", n_estimators),
                        ("max_bins", max_bins)):
        if isinstance(value, bool) or not isinstance(value, int):
            raise TypeError(f"{name} must be an integer")
        if value <= 0:
            raise ValueError(f"{name} must be positive")
    return n_samples * n_features * n_estimators * max_bins


def choose_device(*, requested: RequestedDevice, n_samples: int, n_features: int, n_estimators: int, max_bins: int,
                  hip_available: bool, hip_work_threshold: int = 134_217_728,
                  cpu_available: bool = False) -> DeviceDecision:
    """Choose the backend for one fit and explain the decision."""

    if requested not in {"cpu", "hip", "auto"}:
        raise ValueError("requested device must be 'cpu', 'hip', or 'auto'")
    if not isinstance(hip_available, bool):
        raise TypeError("hip_available must be a bool")
    if isinstance(hip_work_threshold, bool) or not isinstance(hip_work_threshold, int):
        raise TypeError("hip_work_threshold must be an integer")
    if hip_work_threshold <= 0:
        raise ValueError("hip_work_threshold must be positive")
    work = _estimated_histogram_work(n_samples=n_samples, n_features=n_features, n_estimators=n_estimators,
                                     max_bins=max_bins)
    if requested == "cpu":
        return DeviceDecision("cpu", "cpu", "explicit cpu request", work, hip_available)
    if requested == "hip":
        return DeviceDecision("hip", "hip", "explicit hip request", work, hip_available)
    if not hip_available:
        return DeviceDecision("auto", "cpu", "hip unavailable", work, hip_available)
    if work < hip_work_threshold and cpu_available:
        return DeviceDecision("auto", "cpu", "estimated work below hip threshold", work, hip_available)
    if work < hip_work_threshold:
        return DeviceDecision("auto", "hip", "estimated work below hip threshold; no cpu trainer in this package",
                              work, hip_available)
    return DeviceDecision("auto", "hip", "estimated work meets hip threshold", work, hip_available)


def decision_to_dict(decision: DeviceDecision) -> dict[str, object]:
    """A stable, serialisable record of a decision."""

    return {
        "requested": decision.requested,
        "selected": decision.selected,
        "reason": decision.reason,
        "estimated_work": int(decision.estimated_work),
        "hip_available": bool(decision.hip_available),
    }


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
