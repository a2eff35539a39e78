"""Trial grids: :class:`GridSpec`, the size guard and the parametric grids.

``GridSpec`` keeps cuPeriod's contract (src/cuperiod/core/grid.py:104-186, 201-277):
finite ascending values, a frequency or period view, rounding-level steps back
tolerated where concatenated ``np.arange`` segments meet, and ``meta`` for the BLS
segmentation. ``check_grid_size`` refuses a grid above :data:`MAX_GRID_SIZE` before
anything is allocated and warns about large automatic grids (``:417-505``).

Two helpers describe grids *parametrically* so a batch never materialises per-curve
arrays on the host: :func:`linspace_frequencies` reproduces ``np.linspace`` bit for bit
from ``(start, stop, num)``, which lets the device regenerate the same values.
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field
from typing import Any, Literal, Mapping

import numpy as np

from cuperiod_amd.core.errors import GridSizeWarning, GridTooLargeError

GridKind = Literal["frequency", "period"]

#: Hard ceiling on the number of trial samples of one grid.
MAX_GRID_SIZE: int = 250_000_000

#: Automatic grids above this size warn (and are still built).
AUTO_GRID_WARN_SIZE: int = 5_000_000

_SPAN_OUTLIER_RATIO = 10.0
_CHECK_CHUNK = 1 << 20


@dataclass(frozen=True, eq=False)
class GridSpec:
    """A trial grid in the frequency (cycles/day) or period (days) domain."""

    kind: GridKind
    values: np.ndarray
    uniform: bool = False
    meta: Mapping[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        if self.kind not in ("frequency", "period"):
            raise ValueError(
                f"grid kind must be 'frequency' or 'period', got {self.kind!r}"
            )
        values = np.ascontiguousarray(np.asarray(self.values, dtype=np.float64))
        if values.ndim != 1:
            raise ValueError(f"grid values must be 1-D, got shape {values.shape}")
        _validate_grid_values(values, self.kind, bool(self.uniform))
        object.__setattr__(self, "values", values)

    @property
    def size(self) -> int:
        return int(self.values.size)

    @property
    def frequency(self) -> np.ndarray:
        """The grid as ascending frequencies."""
        if self.kind == "frequency":
            return self.values
        return _reciprocal_reversed(self.values)

    @property
    def period(self) -> np.ndarray:
        """The grid as ascending periods."""
        if self.kind == "period":
            return self.values
        return _reciprocal_reversed(self.values)

    def uniform_frequency_params(self) -> tuple[float, float, int]:
        """``(f0, df, nf)`` of a frequency grid."""
        if self.kind != "frequency":
            raise ValueError("uniform_frequency_params requires a frequency grid")
        nf = int(self.values.size)
        if nf == 0:
            return 0.0, 0.0, 0
        f0 = float(self.values[0])
        df = float(self.values[1] - self.values[0]) if nf > 1 else 0.0
        return f0, df, nf


def _reciprocal_reversed(values: np.ndarray) -> np.ndarray:
    """``1/values`` in reverse order as a fresh contiguous array (``1/v[::-1]``)."""
    with np.errstate(divide="ignore"):
        return np.ascontiguousarray(1.0 / values[::-1])


def _check_domain(minimum: float, kind: str) -> None:
    if kind == "frequency" and minimum < 0.0:
        raise ValueError(f"grid frequencies must be >= 0, got minimum {minimum}")
    if kind == "period" and minimum <= 0.0:
        raise ValueError(f"grid periods must be > 0, got minimum {minimum}")


def _raise_not_ascending(detail: str) -> None:
    raise ValueError(
        f"grid values must be ascending ({detail}); sort them (e.g. np.sort), or "
        "build a period grid as kind='period' instead of reversing 1/frequency"
    )


def _validate_grid_values(values: np.ndarray, kind: str, uniform: bool) -> None:
    """Finite, ascending (steps back of at most ``1e-9 * max|v|``), in-domain values;
    an arithmetic progression when ``uniform`` (src/cuperiod/core/grid.py:201-256)."""
    n = int(values.size)
    if n == 0:
        return
    first, last = float(values[0]), float(values[-1])
    if not (math.isfinite(first) and math.isfinite(last)):
        raise ValueError("grid values must be finite")
    _check_domain(first, kind)
    if n == 1:
        return
    if uniform and kind != "frequency":
        raise ValueError(
            "uniform=True is only meaningful for a frequency grid; pass uniform=False"
        )
    df = float(values[1] - values[0])
    if uniform and not df > 0.0:
        raise ValueError("uniform=True needs a strictly increasing frequency grid")
    scale = max(abs(first), abs(last))
    tolerance = 1e-6 * abs(df) + 8.0 * np.finfo(np.float64).eps * scale
    reversal = 1e-9 * scale
    if not last >= first - reversal:
        _raise_not_ascending(f"values[-1]={last!r} < values[0]={first!r}")
    for start in range(0, n - 1, _CHECK_CHUNK):
        chunk = values[start : start + _CHECK_CHUNK + 1]
        step = np.diff(chunk)
        smallest = float(step.min())
        if not smallest >= -reversal:
            if not bool(np.all(np.isfinite(chunk))):
                raise ValueError("grid values must be finite")
            i = start + int(np.argmin(step))
            _raise_not_ascending(
                f"values[{i}]={float(values[i])!r} > "
                f"values[{i + 1}]={float(values[i + 1])!r}"
            )
        if smallest < 0.0:
            _check_domain(float(chunk.min()), kind)
        if uniform:
            deviation = max(float(step.max()) - df, df - smallest)
            if deviation > tolerance:
                raise ValueError(
                    "uniform=True but the values are not an arithmetic progression "
                    f"(a step differs from values[1]-values[0]={df:.6g} by "
                    f"{deviation:.3g}); pass uniform=False"
                )


def _baseline_outlier_ratio(time: np.ndarray) -> tuple[float, float, float]:
    """``(baseline, central span, ratio)`` with the 1% extreme samples trimmed by rank."""
    t = np.sort(np.asarray(time, dtype=np.float64))
    t = t[np.isfinite(t)]
    n = int(t.size)
    if n < 4:
        return (float(t[-1] - t[0]) if n else 0.0), 0.0, 1.0
    k = max(1, math.ceil(0.01 * n))
    if n - 2 * k < 2:
        return float(t[-1] - t[0]), 0.0, 1.0
    baseline = float(t[-1] - t[0])
    central = float(t[n - 1 - k] - t[k])
    if central <= 0.0:
        return baseline, central, (math.inf if baseline > 0.0 else 1.0)
    return baseline, central, baseline / central


def check_grid_size(
    n_points: int,
    *,
    limit: int | None = None,
    automatic: bool = False,
    time: np.ndarray | None = None,
    baseline: float | None = None,
    maximum_frequency: float | None = None,
    context: str = "trial grid",
) -> None:
    """Refuse a grid above ``limit`` (default :data:`MAX_GRID_SIZE`); warn about an
    automatic grid above :data:`AUTO_GRID_WARN_SIZE` (src/cuperiod/core/grid.py:417-505)."""
    hard = MAX_GRID_SIZE if limit is None else int(limit)
    n_points = int(n_points)
    over_hard = n_points > hard
    over_soft = automatic and n_points > AUTO_GRID_WARN_SIZE
    if not (over_hard or over_soft):
        return
    outlier = ""
    if time is not None and np.size(time) > 0:
        span, central, ratio = _baseline_outlier_ratio(time)
        baseline = span if baseline is None else baseline
        if ratio > _SPAN_OUTLIER_RATIO:
            outlier = (
                f" The baseline ({span:.6g} d) is {ratio:.3g}x the span of the central "
                f"98% of the samples ({central:.6g} d), so a few outlying epochs set "
                "the grid size; check for mis-converted timestamps (a 0 or sentinel "
                "time, a JD/MJD mix-up)."
            )
    gib = n_points * 8.0 / 2.0**30
    origin = ""
    if maximum_frequency is not None:
        origin += f" up to {maximum_frequency:.6g} cycles/day"
    if baseline is not None:
        origin += f" over a {baseline:.6g}-day baseline"
    remedy = " Pass an explicit grid= or narrower period bounds"
    if over_hard:
        raise GridTooLargeError(
            f"{context}: {n_points:,} trial samples{origin} ({gib:.3g} GiB per "
            f"float64 array) exceeds the limit of {hard:,}.{outlier}{remedy}; for a "
            "deliberately larger grid raise cuperiod_amd.core.grid.MAX_GRID_SIZE."
        )
    warnings.warn(
        GridSizeWarning(
            f"{context}: the automatic grid has {n_points:,} trial samples{origin} "
            f"({gib:.3g} GiB per float64 array).{outlier}{remedy} to bound it. "
            "Computing it anyway."
        ),
        stacklevel=3,
    )


def log_period_grid(
    *, minimum_period: float, maximum_period: float, n_periods: int
) -> GridSpec:
    """Log-spaced ascending period grid (src/cuperiod/core/grid.py:768-802)."""
    if not (math.isfinite(minimum_period) and math.isfinite(maximum_period)):
        raise ValueError(
            f"minimum_period and maximum_period must be finite, got "
            f"({minimum_period}, {maximum_period})"
        )
    if not 0.0 < minimum_period < maximum_period:
        raise ValueError("require 0 < minimum_period < maximum_period")
    if n_periods < 1:
        raise ValueError("n_periods must be >= 1")
    check_grid_size(n_periods, context="log_period_grid")
    periods = np.geomspace(minimum_period, maximum_period, n_periods)
    return GridSpec(kind="period", values=periods, uniform=False)


def linspace_frequencies(start: float, stop: float, num: int) -> np.ndarray:
    """``np.linspace(start, stop, num)`` (float64, endpoint=True) by its own recipe:
    ``value_i = i*step + start`` with ``step = (stop-start)/(num-1)``, the last sample
    set to ``stop``. The device regenerates a batch's grids with exactly this recipe,
    so a grid travels as three numbers instead of ``num`` values."""
    num = int(num)
    if num <= 0:
        return np.empty(0, dtype=np.float64)
    if num == 1:
        return np.asarray([float(start)], dtype=np.float64)
    div = num - 1
    delta = float(stop) - float(start)
    step = delta / div
    i = np.arange(num, dtype=np.float64)
    if step == 0.0:
        y = (i / div) * delta + float(start)
    else:
        y = i * step + float(start)
    y[-1] = float(stop)
    return y


__all__ = [
    "AUTO_GRID_WARN_SIZE",
    "MAX_GRID_SIZE",
    "GridKind",
    "GridSizeWarning",
    "GridSpec",
    "GridTooLargeError",
    "check_grid_size",
    "linspace_frequencies",
    "log_period_grid",
]
