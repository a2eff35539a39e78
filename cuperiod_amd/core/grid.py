"""Trial grids: :class:`GridSpec`, the size guard and the parametric grid recipe.

A :class:`GridSpec` is an ascending array of trial frequencies (cycles/day) or periods
(days) with a view in the other domain; its fields (``kind``, ``values``, ``uniform``,
``meta``) and views are the interface of cuPeriod's grid object
(src/cuperiod/core/grid.py:104-186), which every method's ``power(grid, ...)`` takes.
Validation happens once, vectorised, when a grid is made.

:func:`check_grid_size` refuses a grid too large to allocate before anything is built and
warns about large automatic grids. :func:`linspace_frequencies` reproduces
``np.linspace`` bit for bit from ``(start, stop, num)``: the engine regenerates a batch's
linspace grids on the device with the same recipe, so such a grid travels as three
numbers.
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field
from typing import Any, Literal, Mapping

import numpy as np

from cuperiod_amd.core.errors import GridSizeWarning, GridTooLargeError

GridKind = Literal["frequency", "period"]

#: Largest number of trial values one grid may hold.
MAX_GRID_SIZE: int = 250_000_000

#: Automatic grids larger than this are built, with a warning.
AUTO_GRID_WARN_SIZE: int = 5_000_000

# Steps back tolerated between neighbours, relative to the largest magnitude: where two
# np.arange segments are concatenated their last and first values may cross by rounding.
_STEP_BACK = 1e-9


def _grid_problem(values: np.ndarray, kind: str, uniform: bool) -> str | None:
    """What is wrong with ``values`` as a grid of ``kind``, or None."""
    if kind not in ("frequency", "period"):
        return f"kind must be 'frequency' or 'period', not {kind!r}"
    if values.ndim != 1:
        return f"values must be one-dimensional, not of shape {values.shape}"
    if values.size == 0:
        return None
    if not np.isfinite(values).all():
        return "values must be finite"
    low = float(values.min())
    if kind == "period" and not low > 0.0:
        return f"periods must be > 0 (smallest {low!r})"
    if kind == "frequency" and low < 0.0:
        return f"frequencies must be >= 0 (smallest {low!r})"
    if values.size == 1:
        return None
    steps = np.diff(values)
    scale = float(np.abs(values).max())
    worst = int(np.argmin(steps))
    if steps[worst] < -_STEP_BACK * scale:
        return (f"values must be ascending: values[{worst}] = {values[worst]!r} comes before "
                f"{values[worst + 1]!r} (sort them, or give periods as kind='period')")
    if uniform:
        if kind != "frequency":
            return "uniform=True describes a frequency grid only"
        step = float(steps[0])
        spread = float(np.abs(steps - step).max())
        if not step > 0.0 or spread > 1e-6 * step + 8.0 * np.finfo(np.float64).eps * scale:
            return "uniform=True, but the values are not evenly spaced"
    return None


@dataclass(frozen=True, eq=False)
class GridSpec:
    """A trial grid: ascending frequencies (cycles/day) or periods (days)."""

    kind: GridKind
    values: np.ndarray
    uniform: bool = False
    meta: Mapping[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        values = np.ascontiguousarray(self.values, dtype=np.float64)
        problem = _grid_problem(values, self.kind, bool(self.uniform))
        if problem is not None:
            raise ValueError(f"GridSpec: {problem}")
        object.__setattr__(self, "values", values)

    @property
    def size(self) -> int:
        return int(self.values.size)

    def _other_domain(self) -> np.ndarray:
        # reciprocals, reversed so the other view ascends too
        with np.errstate(divide="ignore"):
            return np.ascontiguousarray(1.0 / self.values[::-1])

    @property
    def frequency(self) -> np.ndarray:
        """Ascending frequencies."""
        return self.values if self.kind == "frequency" else self._other_domain()

    @property
    def period(self) -> np.ndarray:
        """Ascending periods."""
        return self.values if self.kind == "period" else self._other_domain()

    def uniform_frequency_params(self) -> tuple[float, float, int]:
        """``(first frequency, step, count)`` of a frequency grid."""
        if self.kind != "frequency":
            raise ValueError("uniform_frequency_params needs a frequency grid")
        n = self.size
        first = float(self.values[0]) if n else 0.0
        step = float(self.values[1] - self.values[0]) if n > 1 else 0.0
        return first, step, n


def _epoch_hint(time: np.ndarray | None) -> str:
    """A remark when a few outlying epochs stretch the baseline (ten times the span of
    the central 98% of the samples, or more)."""
    if time is None:
        return ""
    t = np.asarray(time, dtype=np.float64)
    t = t[np.isfinite(t)]
    if t.size < 4:
        return ""
    full = float(t.max() - t.min())
    lo, hi = np.quantile(t, [0.01, 0.99])
    central = float(hi - lo)
    if full <= 0.0 or (central > 0.0 and full < 10.0 * central):
        return ""
    return (f" The time baseline ({full:.6g} d) is set by a few outlying epochs (the central "
            f"98% span {central:.6g} d): check for a zero or sentinel time, or a mix of JD "
            "and MJD.")


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
    """Raise :class:`GridTooLargeError` for a grid of more than ``limit`` values
    (:data:`MAX_GRID_SIZE` by default), before anything is allocated; warn with
    :class:`GridSizeWarning` for an automatic grid above :data:`AUTO_GRID_WARN_SIZE`
    (the contract of src/cuperiod/core/grid.py:417-505)."""
    cap = MAX_GRID_SIZE if limit is None else int(limit)
    n = int(n_points)
    too_big = n > cap
    if not too_big and not (automatic and n > AUTO_GRID_WARN_SIZE):
        return
    what = f"{context}: {n:,} trial values ({n * 8 / 2**30:.3g} GiB as float64)"
    if maximum_frequency is not None:
        what += f", up to {maximum_frequency:.6g} cycles/day"
    if baseline is not None:
        what += f", for a {baseline:.6g}-day baseline"
    what += "." + _epoch_hint(time)
    if too_big:
        raise GridTooLargeError(
            f"{what} The limit is {cap:,} (cuperiod_amd.core.grid.MAX_GRID_SIZE); pass an "
            "explicit grid or narrower period bounds.")
    warnings.warn(GridSizeWarning(f"{what} Bound it with an explicit grid or narrower period "
                                  "bounds; it is built anyway."), stacklevel=3)


def log_period_grid(*, minimum_period: float, maximum_period: float,
                    n_periods: int) -> GridSpec:
    """``n_periods`` periods spaced evenly in log between the two bounds."""
    lo, hi, n = float(minimum_period), float(maximum_period), int(n_periods)
    if not (math.isfinite(lo) and math.isfinite(hi) and 0.0 < lo < hi):
        raise ValueError(f"log_period_grid needs finite 0 < minimum_period < maximum_period, "
                         f"got {lo!r}, {hi!r}")
    if n < 1:
        raise ValueError(f"log_period_grid needs n_periods >= 1, got {n}")
    check_grid_size(n, context="log_period_grid")
    return GridSpec(kind="period", values=np.exp(np.linspace(math.log(lo), math.log(hi), n)))


def linspace_frequencies(start: float, stop: float, num: int) -> np.ndarray:
    """``np.linspace(start, stop, num)`` (float64, endpoint included) by its own recipe:
    ``value_i = i*step + start`` with ``step = (stop-start)/(num-1)`` and the last value
    set to ``stop`` (``(i/(num-1))*(stop-start) + start`` when the step underflows)."""
    num = int(num)
    if num <= 0:
        return np.empty(0, dtype=np.float64)
    if num == 1:
        return np.asarray([float(start)], dtype=np.float64)
    div = num - 1
    delta = float(stop) - float(start)
    step = delta / div
    i = np.arange(num, dtype=np.float64)
    values = i * step + float(start) if step != 0.0 else (i / div) * delta + float(start)
    values[-1] = float(stop)
    return values


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
