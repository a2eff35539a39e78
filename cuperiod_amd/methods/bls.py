"""Box Least Squares method on the MI355X engine.

The search runs in log-spaced period segments so the box durations track the period
(cuPeriod src/cuperiod/methods/bls.py:39-190): the default grid is a set of segments
``[p_lo, factor*p_lo)``, each uniform in frequency with step ``grid_duration_frac /
(grid_oversample * baseline)`` and durations ``geomspace(max(min_frac*p_lo,
min_duration_days), max_frac*p_lo, n_durations)``; a custom grid is split into the same
log segments from its shortest period. Every segment of a light curve is searched in
one call of the HIP engine; the segment outputs are assembled into a
:class:`~cuperiod_amd.core.result.Periodogram` with the signal detection efficiency
(SDE) as ``_assemble`` does (``:274-301``).
"""

from __future__ import annotations

import math
from typing import Any, ClassVar

import numpy as np

from cuperiod_amd.core.columns import Domain
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.core.errors import InsufficientDataError
from cuperiod_amd.core.grid import GridSpec, check_grid_size
from cuperiod_amd.core.lightcurve import LightCurve
from cuperiod_amd.core.result import Periodogram
from cuperiod_amd.methods._bls_core import HIP_BACKEND, bls_power_segments
from cuperiod_amd.methods.base import PeriodogramMethod, register, searchable_points

#: Per-period fields every search reports.
_SEGMENT_FIELDS = ("period", "power", "depth", "depth_snr", "duration", "transit_time")

Segment = tuple[np.ndarray, np.ndarray]


def _segment_durations(p_lo: float, settings: BLSSettings) -> np.ndarray:
    """Box durations (days) of the segment starting at ``p_lo``."""
    d_hi = settings.duration_max_frac * p_lo
    d_lo = max(settings.duration_min_frac * p_lo, settings.min_duration_days)
    if d_lo >= d_hi:
        return np.asarray([d_hi], dtype=np.float64)
    return np.asarray(np.geomspace(d_lo, d_hi, settings.n_durations), dtype=np.float64)


def _max_period(baseline: float, settings: BLSSettings) -> float:
    """Longest searchable period: at least ``min_transits`` transits must fit."""
    return min(settings.max_period_days, baseline / settings.min_transits)


def _segment_bounds(baseline: float, settings: BLSSettings) -> list[tuple[float, float]]:
    """``[(p_lo, p_hi), ...]`` of the default log segments, ascending: edges from
    ``min_period_days``, each the previous one times ``segment_factor``, the last one
    clipped to the longest searchable period."""
    top = _max_period(baseline, settings)
    edges = [settings.min_period_days]
    while edges[-1] < top:
        edges.append(min(edges[-1] * settings.segment_factor, top))
    return list(zip(edges[:-1], edges[1:]))


def _segment_periods(
    baseline: float,
    settings: BLSSettings,
    *,
    time: np.ndarray | None = None,
    context: str = "BLS default grid",
) -> list[tuple[float, np.ndarray]]:
    """``[(p_lo, periods), ...]`` of the non-empty default segments, ascending; the
    total size is checked before allocation. Each segment's frequencies are
    ``np.arange(1/p_hi, 1/p_lo, df)`` without a sample at or above ``1/p_lo``."""
    bounds = _segment_bounds(baseline, settings)
    if not bounds:
        return []
    df = settings.grid_duration_frac / (settings.grid_oversample * baseline)
    total = sum(max(0, math.ceil((1.0 / p_lo - 1.0 / p_hi) / df)) for p_lo, p_hi in bounds)
    check_grid_size(total, automatic=True, time=time, baseline=baseline, context=context)
    def periods_of(p_lo: float, p_hi: float) -> np.ndarray:
        f_top = 1.0 / p_lo
        freq = np.arange(1.0 / p_hi, f_top, df)
        return np.asarray(1.0 / freq[freq < f_top][::-1], dtype=np.float64)

    segments = ((p_lo, periods_of(p_lo, p_hi)) for p_lo, p_hi in bounds)
    return [(p_lo, periods) for p_lo, periods in segments if periods.size]


def _segment_grids(
    baseline: float,
    settings: BLSSettings,
    *,
    time: np.ndarray | None = None,
    context: str = "BLS default grid",
) -> list[Segment]:
    return [
        (periods, _segment_durations(p_lo, settings))
        for p_lo, periods in _segment_periods(baseline, settings, time=time, context=context)
    ]


def _log_segment_index(periods: np.ndarray, settings: BLSSettings) -> np.ndarray:
    """Segment index ``floor(log(P/p0)/log(factor) + 1e-9)`` of every period."""
    p0 = float(periods.min())
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.floor(
            np.log(periods / p0) / math.log(settings.segment_factor) + 1e-9
        ).astype(np.int64)


def _log_segments(periods: np.ndarray, settings: BLSSettings) -> list[Segment]:
    """Split an arbitrary period grid into log segments from its shortest period
    ``p0``: segment ``k`` holds ``[p0 f^k, p0 f^(k+1))`` with the durations of a
    default segment starting at ``p0 f^k``; periods keep their order."""
    p0 = float(periods.min())
    index = _log_segment_index(periods, settings)
    return [(np.ascontiguousarray(periods[index == k]),
             _segment_durations(p0 * settings.segment_factor**k, settings))
            for k in np.unique(index).tolist()]


def _grid_segments(grid: GridSpec, settings: BLSSettings) -> list[Segment]:
    """The ``(periods, durations)`` segments that search exactly ``grid``'s periods:
    the recorded default segmentation when ``grid.meta`` holds a consistent one,
    else :func:`_log_segments`."""
    periods = np.asarray(grid.period, dtype=np.float64)
    if periods.size == 0:
        return []
    if not np.all(np.isfinite(periods)):
        raise ValueError("BLS: trial periods must be finite; drop frequency 0 from the grid")
    starts, sizes = grid.meta.get("segment_starts"), grid.meta.get("segment_sizes")
    if starts is None or sizes is None or len(starts) != len(sizes):
        return _log_segments(periods, settings)
    ends = np.cumsum([int(n) for n in sizes], dtype=np.int64)
    if ends.size == 0 or int(ends[-1]) != periods.size:
        return _log_segments(periods, settings)
    return [(periods[int(stop) - int(n):int(stop)], _segment_durations(float(p_lo), settings))
            for p_lo, n, stop in zip(starts, sizes, ends)]


def _segments_power(
    backend: str,
    jd: np.ndarray,
    flux: np.ndarray,
    err: np.ndarray,
    segments: list[Segment],
    settings: BLSSettings,
) -> list[dict[str, np.ndarray]]:
    """Every segment's per-period maxima of one light curve, in one engine call."""
    powers = bls_power_segments(
        jd, flux, err, segments, settings.bins_per_duration,
        objective=settings.objective, backend=backend, batch=settings.batch_periods,
        precision=settings.precision, device_cache={}, accumulation=settings.accumulation,
    )
    return [{name: getattr(p, name) for name in _SEGMENT_FIELDS} for p in powers]


def sde_of(power: np.ndarray) -> np.ndarray:
    """Signal detection efficiency ``(p - mean)/std`` (0 where ``std == 0``)."""
    mean, std = float(power.mean()), float(power.std())
    return (power - mean) / std if std > 0.0 else np.zeros_like(power)


#: Periodogram extras and the engine output each one is read from.
_EXTRA_SOURCES = (("depth", "depth"), ("depth_snr", "depth_snr"), ("duration", "duration"),
                  ("t0", "transit_time"))


def _assemble(
    chunks: dict[str, list[np.ndarray]], n: int, baseline: float, backend: str, meta: Any
) -> Periodogram:
    """Concatenate segment outputs into one periodogram: power with NaN/+-inf set to
    0, SDE, ``frequency = 1/period``, ascending-frequency order."""
    joined = {key: np.concatenate(parts) for key, parts in chunks.items()}
    power = np.nan_to_num(joined["power"], nan=0.0, posinf=0.0, neginf=0.0)
    extras = {name: joined[source] for name, source in _EXTRA_SOURCES}
    extras["sde"] = sde_of(power)
    return Periodogram.from_spectrum(
        method="BLS", backend=backend, objective_sense="max",
        frequency=1.0 / joined["period"], power=power, extras=extras,
        n_samples=n, baseline=baseline, meta=meta,
    )


class BLSMethod(PeriodogramMethod):
    """Box Least Squares on the MI355X HIP engine (backend ``"hip"``)."""

    name: ClassVar[str] = "BLS"
    objective_sense: ClassVar[str] = "max"
    supports_multiband: ClassVar[bool] = True
    natural_domain: ClassVar[Domain] = Domain.FLUX
    settings_cls: ClassVar[type] = BLSSettings
    cpu_backend: ClassVar[str | None] = None
    fast_cpu_backend: ClassVar[str | None] = None
    gpu_backend: ClassVar[str | None] = HIP_BACKEND
    portable_gpu_backend: ClassVar[str | None] = None
    all_backends: ClassVar[tuple[str, ...]] = (HIP_BACKEND,)

    def default_grid(self, lc: LightCurve, settings: BLSSettings) -> GridSpec:
        """Log period segments, each uniform in frequency, up to ``min(max_period_days,
        baseline/min_transits)``; ``meta`` records the segmentation."""
        finite = lc.finite()
        segments = _segment_periods(finite.baseline, settings, time=finite.time)
        if not segments:
            raise InsufficientDataError(
                "BLS: baseline too short for the configured period range")
        starts = tuple(float(p_lo) for p_lo, _ in segments)
        sizes = tuple(int(periods.size) for _, periods in segments)
        return GridSpec(kind="period", uniform=False,
                        values=np.concatenate([periods for _, periods in segments]),
                        meta={"segmented": True, "segment_starts": starts,
                              "segment_sizes": sizes})

    def power(
        self,
        grid: GridSpec,
        lc: LightCurve,
        settings: BLSSettings,
        backend: str,
        engine: object | None = None,
    ) -> Periodogram:
        finite = searchable_points(lc, settings.min_detections, "BLS")
        segments = _grid_segments(grid, settings)
        if not segments:
            raise InsufficientDataError("BLS: empty trial grid")
        err = finite.error if finite.error is not None else np.ones_like(finite.value)
        found = _segments_power(backend, finite.time, finite.value, err, segments, settings)
        chunks = {name: [seg[name] for seg in found] for name in _SEGMENT_FIELDS}
        return _assemble(chunks, finite.n, finite.baseline, backend, finite.meta)

    def multiband_power(
        self,
        grid: GridSpec,
        mblc: Any,
        settings: BLSSettings,
        backend: str,
        engine: object | None = None,
    ) -> Periodogram:
        """Multi-band BLS (src/cuperiod/methods/bls.py:397-411): the independent stack
        or the joint-ephemeris search, as ``settings.mb_ephemeris`` says."""
        from cuperiod_amd.methods.bls_mb import bls_multiband_power

        return bls_multiband_power(grid, mblc, settings, backend, self)

    def estimate_device_bytes(self, n_points: int) -> int:
        return 128 * 1024**2 + n_points * 8 * 10


register(BLSMethod())

__all__ = ["BLSMethod", "sde_of"]
