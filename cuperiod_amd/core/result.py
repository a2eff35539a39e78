"""Result objects: :class:`Peak`, :class:`Periodogram` and :class:`MultiResult`.

A :class:`Periodogram` is one method's spectrum of one light curve, stored in ascending
frequency, with the fields of cuPeriod's result (src/cuperiod/core/result.py:81-140) and
its peak reporting: candidates are the interior local maxima plus the boundary samples
that beat their neighbour, ranked by score with a minimum frequency separation, or
alias-diverse for the box searches (the rule of src/cuperiod/core/result.py:285-438 and
src/cuperiod/core/peaks.py:80-231, implemented in :mod:`cuperiod_amd.core.peaks`). That
rule decides which period is reported, so it, including how equal scores are ordered, is
kept as it is.
"""

from __future__ import annotations

import dataclasses
from collections.abc import Mapping as MappingABC
from dataclasses import dataclass, field
from typing import Any, Iterator, Literal, Mapping

import numpy as np

from cuperiod_amd.core.peaks import (
    peak_preserving_downsample,
    select_diverse_peaks,
    select_top_peaks,
)

ObjectiveSense = Literal["max", "min"]

DEFAULT_N_PEAKS = 10
DEFAULT_SEPARATION_RAYLEIGH = 3.0
DEFAULT_HARMONIC_MAX = 8
DEFAULT_DOWNSAMPLE_POINTS = 2000

#: Peak-selection parameter -> the settings field that supplies its default.
PEAK_DEFAULT_FIELDS: Mapping[str, str] = {
    "n": "n_peaks",
    "min_separation_rayleigh": "peak_separation_rayleigh",
    "alias_tolerance": "alias_freq_tolerance",
    "harmonic_max": "harmonic_max",
    "n_points": "downsample_points",
}


def peak_defaults_from_settings(settings: object | None) -> dict[str, Any]:
    """The peak-selection defaults a settings object supplies (fields it does not have,
    or leaves None, supply none)."""
    if settings is None:
        return {}
    found = {p: getattr(settings, f, None) for p, f in PEAK_DEFAULT_FIELDS.items()}
    return {p: v for p, v in found.items() if v is not None}


@dataclass(frozen=True)
class Peak:
    """One reported peak: period, frequency, power, 1-based rank and the method's scalars
    at it (BLS: depth, depth_snr, duration, t0, sde)."""

    period: float
    frequency: float
    power: float
    rank: int
    extra: Mapping[str, float] = field(default_factory=dict, hash=False)

    def to_dict(self) -> dict[str, float | int]:
        return {"period": self.period, "frequency": self.frequency, "power": self.power,
                "rank": self.rank, **self.extra}


def _sense_checked(sense: str) -> str:
    if sense not in ("max", "min"):
        raise ValueError(f"objective_sense is 'max' or 'min', not {sense!r}")
    return sense


@dataclass(frozen=True, eq=False)
class Periodogram:
    """One method's spectrum of one light curve, in ascending frequency
    (``period = 1/frequency``); ``extras`` are per-sample arrays of method scalars."""

    method: str
    backend: str
    frequency: np.ndarray
    period: np.ndarray
    power: np.ndarray
    objective_sense: ObjectiveSense
    n_samples: int
    baseline: float
    extras: Mapping[str, np.ndarray] = field(default_factory=dict)
    meta: Mapping[str, Any] = field(default_factory=dict)
    peak_defaults: Mapping[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        _sense_checked(self.objective_sense)

    @classmethod
    def from_spectrum(cls, *, method: str, backend: str, frequency: np.ndarray,
                      power: np.ndarray, objective_sense: ObjectiveSense, n_samples: int,
                      baseline: float, extras: Mapping[str, np.ndarray] | None = None,
                      meta: Mapping[str, Any] | None = None,
                      peak_defaults: Mapping[str, Any] | None = None,
                      order: np.ndarray | None = None) -> "Periodogram":
        """A periodogram from samples in any order: every array is put in ascending
        frequency by a stable sort (``order``, when given, must be that permutation: a
        grid known to be descending in frequency passes its reversal)."""
        _sense_checked(objective_sense)
        freq = np.asarray(frequency, dtype=np.float64)
        if freq.ndim != 1:
            raise ValueError(f"from_spectrum: frequency must be one-dimensional, not of "
                             f"shape {freq.shape}")
        columns = {"power": power, **dict(extras or {})}
        arrays = {}
        for name, data in columns.items():
            arr = np.asarray(data, dtype=np.float64)
            if arr.shape != freq.shape:
                raise ValueError(f"from_spectrum: {name} has shape {arr.shape}, the "
                                 f"frequencies {freq.shape}")
            arrays[name] = arr
        perm = np.argsort(freq, kind="stable") if order is None else np.asarray(order)
        ordered = {k: np.ascontiguousarray(v[perm]) for k, v in arrays.items()}
        freq = np.ascontiguousarray(freq[perm])
        with np.errstate(divide="ignore"):
            period = 1.0 / freq
        power_sorted = ordered.pop("power")
        return cls(method=method, backend=backend, frequency=freq, period=period,
                   power=power_sorted, objective_sense=objective_sense, n_samples=n_samples,
                   baseline=baseline, extras=ordered, meta=dict(meta or {}),
                   peak_defaults=dict(peak_defaults or {}))

    def with_peak_defaults(self, defaults: Mapping[str, Any] | object) -> "Periodogram":
        """A copy whose peak-selection defaults are updated from a mapping (keys of
        :data:`PEAK_DEFAULT_FIELDS`) or a settings object."""
        update = (dict(defaults) if isinstance(defaults, MappingABC)
                  else peak_defaults_from_settings(defaults))
        extra = set(update) - set(PEAK_DEFAULT_FIELDS)
        if extra:
            raise ValueError(f"with_peak_defaults: {sorted(extra)} are not peak settings "
                             f"(known: {sorted(PEAK_DEFAULT_FIELDS)})")
        return dataclasses.replace(self, peak_defaults={**self.peak_defaults, **update})

    def _default(self, name: str, fallback: Any) -> Any:
        value = self.peak_defaults.get(name)
        return fallback if value is None else value

    @property
    def size(self) -> int:
        return int(self.frequency.size)

    def _score(self) -> np.ndarray:
        """Larger is more significant."""
        return self.power if self.objective_sense == "max" else -self.power

    def _peak_candidates(self, score: np.ndarray) -> np.ndarray:
        """Candidate samples, in the order the selection sees them: interior local
        maxima (above the left neighbour, at least the right one: a flat top is reported
        at its left edge) ascending, then sample 0 if it is at least sample 1, then the
        last sample if it beats its neighbour. Samples with a non-finite score or period
        never qualify (they compare as -inf); with no candidate, the best finite sample."""
        usable = np.isfinite(score) & np.isfinite(self.period)
        s = np.where(usable, score, -np.inf)
        n = s.size
        if n >= 3:
            mid = np.flatnonzero((s[1:-1] > s[:-2]) & (s[1:-1] >= s[2:])) + 1
        else:
            mid = np.zeros(0, dtype=np.int64)
        ends = []
        if n == 1:
            ends = [0]
        elif n >= 2:
            if s[0] >= s[1]:
                ends.append(0)
            if s[-1] > s[-2]:
                ends.append(n - 1)
        cand = np.concatenate([mid, np.asarray(ends, dtype=np.int64)]).astype(np.int64)
        cand = cand[usable[cand]]
        if cand.size == 0 and usable.any():
            ok = np.flatnonzero(usable)
            cand = ok[[int(np.argmax(score[ok]))]]
        return cand

    def select_peaks(self, frequency: np.ndarray, score: np.ndarray, candidates: np.ndarray,
                     count: int, *, min_separation_rayleigh: float | None = None,
                     alias_diverse: bool = False, harmonic_max: int | None = None,
                     alias_tolerance: float | None = None,
                     examined: list | None = None) -> np.ndarray:
        """Indices of the ``count`` peaks :meth:`best_periods` reports, chosen among
        ``candidates`` with this periodogram's baseline and defaults (``examined``
        receives the number of candidates visited)."""
        rayleigh = float(min_separation_rayleigh if min_separation_rayleigh is not None
                         else self._default("min_separation_rayleigh",
                                            DEFAULT_SEPARATION_RAYLEIGH))
        resolved = self.baseline > 0.0
        separation = rayleigh / self.baseline if resolved else 0.0
        if not alias_diverse:
            return select_top_peaks(frequency, score, candidates, count, separation,
                                    examined=examined)
        order = int(harmonic_max if harmonic_max is not None
                    else self._default("harmonic_max", DEFAULT_HARMONIC_MAX))
        tol = alias_tolerance if alias_tolerance is not None else self._default(
            "alias_tolerance", None)
        if tol is None:
            tol = min(separation, 1.0 / self.baseline) if resolved else 0.0
        return select_diverse_peaks(frequency, score, candidates, count,
                                    max(float(tol), 1e-12), order,
                                    min_separation=max(separation, 1e-12), examined=examined)

    def best_periods(self, n: int | None = None, *,
                     min_separation_rayleigh: float | None = None,
                     alias_diverse: bool = False, harmonic_max: int | None = None,
                     alias_tolerance: float | None = None) -> list[Peak]:
        """The ``n`` most significant peaks, most significant first."""
        count = int(self._default("n", DEFAULT_N_PEAKS) if n is None else n)
        if count <= 0 or self.size == 0:
            return []
        score = self._score()
        cand = self._peak_candidates(score)
        if cand.size == 0:
            return []
        picks = self.select_peaks(self.frequency, score, cand, count,
                                  min_separation_rayleigh=min_separation_rayleigh,
                                  alias_diverse=alias_diverse, harmonic_max=harmonic_max,
                                  alias_tolerance=alias_tolerance)
        return [Peak(period=float(self.period[i]), frequency=float(self.frequency[i]),
                     power=float(self.power[i]), rank=r,
                     extra={k: float(v[i]) for k, v in self.extras.items()})
                for r, i in enumerate(picks.tolist(), 1)]

    def best_period(self) -> float:
        """The most significant period in days (NaN when nothing is finite)."""
        top = self.best_periods(1)
        return top[0].period if top else float("nan")

    def downsample(self, n_points: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        """``(frequency, power)`` thinned to ``n_points`` samples that keep each index
        bin's most significant sample (float32)."""
        count = int(self._default("n_points", DEFAULT_DOWNSAMPLE_POINTS)
                    if n_points is None else n_points)
        return peak_preserving_downsample(self.frequency, self.power, count,
                                          sense=self.objective_sense)


def _method_key(name: object) -> str:
    return "".join(ch for ch in str(name) if ch.isalnum()).upper()


class MultiResult(MappingABC):
    """Periodograms of one light curve by several methods, looked up by method name
    (case and separators ignored: ``"bls"`` finds ``"BLS"``)."""

    def __init__(self, results: Mapping[str, Periodogram]) -> None:
        self.results = dict(results)

    def __getitem__(self, method: str) -> Periodogram:
        for name, pg in self.results.items():
            if name == method or (isinstance(method, str)
                                  and _method_key(name) == _method_key(method)):
                return pg
        raise KeyError(method)

    def __iter__(self) -> Iterator[str]:
        return iter(self.results)

    def __len__(self) -> int:
        return len(self.results)

    def best_periods(self, n: int | None = None, **kwargs: Any) -> dict[str, list[Peak]]:
        return {name: pg.best_periods(n, **kwargs) for name, pg in self.results.items()}


__all__ = [
    "DEFAULT_DOWNSAMPLE_POINTS",
    "DEFAULT_HARMONIC_MAX",
    "DEFAULT_N_PEAKS",
    "DEFAULT_SEPARATION_RAYLEIGH",
    "PEAK_DEFAULT_FIELDS",
    "MultiResult",
    "ObjectiveSense",
    "Peak",
    "Periodogram",
    "peak_defaults_from_settings",
]
