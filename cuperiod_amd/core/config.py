"""BLS and TLS settings: the fields, defaults, bounds and environment overrides of
cuPeriod's ``BLSSettings`` (src/cuperiod/core/config.py:218-302, device/precision from
``:59-101``) and ``TLSSettings`` (``:605-702``).

Every field can be overridden from the environment as ``CUPERIOD_BLS_<FIELD>`` /
``CUPERIOD_TLS_<FIELD>`` (for example ``CUPERIOD_BLS_N_DURATIONS=6``), as the
reference's pydantic-settings models allow; an explicit constructor argument wins over
the environment.

One field is new: ``accumulation`` selects the engine's precision (``"fixed64"``: FP64
phase fold and exact 64-bit fixed-point bins, the default; ``"kahan32"``: FP64 phase fold,
FP32 bins with TwoSum-compensated atomic adds and double-float prefix sums; ``"fp32"``:
all FP32 -- phase fold, bins and prefix sums -- the precision ablation). Both FP32 modes
evaluate boxes in FP32. ``backend`` accepts ``"hip"``, this engine's backend name.
"""

from __future__ import annotations

import json
import os
from typing import Any, Literal

from pydantic import BaseModel, ConfigDict, Field, model_validator

ENV_PREFIX = "CUPERIOD_BLS_"
TLS_ENV_PREFIX = "CUPERIOD_TLS_"


def _from_env(cls: type[BaseModel], data: Any, prefix: str) -> Any:
    """``data`` with the fields it does not set filled from ``<prefix><FIELD>``."""
    data = dict(data) if isinstance(data, dict) else data
    if not isinstance(data, dict):
        return data
    for name in cls.model_fields:
        if name in data:
            continue
        raw = None
        for key, value in os.environ.items():
            if key.upper() == prefix + name.upper():
                raw = value
        if raw is None:
            continue
        try:
            data[name] = json.loads(raw)
        except (ValueError, TypeError):
            data[name] = raw
    return data


def _require_lt(lo: float | None, hi: float | None, lo_name: str, hi_name: str) -> None:
    if lo is not None and hi is not None and lo >= hi:
        raise ValueError(f"{lo_name} ({lo}) must be < {hi_name} ({hi})")


class BLSSettings(BaseModel):
    """Settings of the box least squares search."""

    model_config = ConfigDict(extra="forbid", validate_assignment=True, frozen=False)

    @model_validator(mode="before")
    @classmethod
    def _from_environment(cls, data: Any) -> Any:
        """Fill fields not passed explicitly from ``CUPERIOD_BLS_<FIELD>``."""
        return _from_env(cls, data, ENV_PREFIX)

    @model_validator(mode="after")
    def _check_bounds(self) -> "BLSSettings":
        _require_lt(
            self.min_period_days, self.max_period_days,
            "min_period_days", "max_period_days",
        )
        _require_lt(
            self.duration_min_frac, self.duration_max_frac,
            "duration_min_frac", "duration_max_frac",
        )
        return self

    device: Literal["auto", "cpu", "cuda", "mps", "xpu"] = Field(default="auto")
    precision: Literal["auto", "float64", "float32"] = Field(default="auto")
    min_period_days: float = Field(default=0.2, gt=0.0)
    max_period_days: float = Field(default=100.0, gt=0.0)
    min_transits: int = Field(default=2, ge=1)
    duration_min_frac: float = Field(default=0.01, gt=0.0, lt=1.0)
    duration_max_frac: float = Field(default=0.12, gt=0.0, lt=1.0)
    n_durations: int = Field(default=8, ge=1)
    min_duration_days: float = Field(default=0.01, gt=0.0)
    segment_factor: float = Field(default=2.0, gt=1.0)
    grid_duration_frac: float = Field(default=0.05, gt=0.0)
    grid_oversample: int = Field(default=2, ge=1)
    bins_per_duration: int = Field(default=10, ge=1)
    objective: Literal["snr", "likelihood"] = Field(default="snr")
    mb_ephemeris: Literal["independent", "shared"] = Field(default="independent")
    n_peaks: int = Field(default=10, ge=1)
    peak_separation_rayleigh: float = Field(default=3.0, gt=0.0)
    alias_freq_tolerance: float = Field(default=0.0035, ge=0.0)
    harmonic_max: int = Field(default=8, ge=1)
    min_detections: int = Field(default=20, ge=3)
    backend: Literal[
        "auto", "cpu", "gpu", "numba", "numpy", "astropy", "cupy", "torch", "hip"
    ] = Field(default="auto")
    batch_periods: int = Field(default=2048, ge=1)
    downsample_points: int = Field(default=2000, ge=2)
    accumulation: Literal["fixed64", "kahan32", "fp32"] = Field(default="fixed64")


class TLSSettings(BaseModel):
    """Settings of the transit least squares (TLS) matched filter."""

    model_config = ConfigDict(extra="forbid", validate_assignment=True, frozen=False)

    @model_validator(mode="before")
    @classmethod
    def _from_environment(cls, data: Any) -> Any:
        """Fill fields not passed explicitly from ``CUPERIOD_TLS_<FIELD>``."""
        return _from_env(cls, data, TLS_ENV_PREFIX)

    @model_validator(mode="after")
    def _check_bounds(self) -> "TLSSettings":
        _require_lt(
            self.min_period_days, self.max_period_days,
            "min_period_days", "max_period_days",
        )
        _require_lt(
            self.duration_min_frac, self.duration_max_frac,
            "duration_min_frac", "duration_max_frac",
        )
        # the quadratic law must stay non-negative at the limb (mu = 0)
        if self.limb_dark_u1 + self.limb_dark_u2 > 1.0:
            raise ValueError(
                f"limb_dark_u1 + limb_dark_u2 ({self.limb_dark_u1} + {self.limb_dark_u2}) "
                "must be <= 1 (non-negative limb intensity)"
            )
        return self

    device: Literal["auto", "cpu", "cuda", "mps", "xpu"] = Field(default="auto")
    precision: Literal["auto", "float64", "float32"] = Field(default="auto")
    min_period_days: float = Field(default=0.5, gt=0.0)
    max_period_days: float = Field(default=100.0, gt=0.0)
    min_transits: int = Field(default=2, ge=1)
    oversample: int = Field(default=2, ge=1)
    grid_duration_frac: float = Field(default=0.05, gt=0.0)
    n_phase_bins: int = Field(default=256, ge=16)
    duration_min_frac: float = Field(default=0.01, gt=0.0, lt=1.0)
    duration_max_frac: float = Field(default=0.10, gt=0.0, lt=1.0)
    n_durations: int = Field(default=5, ge=1)
    limb_dark_u1: float = Field(default=0.4, allow_inf_nan=False)
    limb_dark_u2: float = Field(default=0.3, allow_inf_nan=False)
    n_peaks: int = Field(default=10, ge=1)
    peak_separation_rayleigh: float = Field(default=3.0, gt=0.0)
    alias_freq_tolerance: float = Field(default=0.0035, ge=0.0)
    harmonic_max: int = Field(default=8, ge=1)
    min_detections: int = Field(default=20, ge=3)
    backend: Literal["auto", "cpu", "gpu", "numba", "numpy", "cupy", "torch", "hip"] = Field(
        default="auto"
    )
    period_batch: int = Field(default=256, ge=1)
    downsample_points: int = Field(default=2000, ge=2)


__all__ = ["BLSSettings", "ENV_PREFIX", "TLSSettings", "TLS_ENV_PREFIX"]
