"""Result objects: :class:`Peak`, :class:`Periodogram`, :class:`MultiResult`.

Same fields and behaviour as cuPeriod's (src/cuperiod/core/result.py:81-558): a
periodogram is stored in ascending frequency (stable sort), knows whether its statistic
is maximised or minimised, and ranks peaks with the candidate rule and selectors of
:mod:`cuperiod_amd.core.peaks`. ``peak_defaults`` carries the settings' peak fields.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Any, Iterator, Literal, Mapping

import numpy as np

from cuperiod_amd.core.peaks import (
    _interior_maxima,
    _sum_is_finite,
    peak_preserving_downsample,
    select_diverse_peaks,
    select_top_peaks,
)

ObjectiveSense = Literal["max", "min"]

DEFAULT_N_PEAKS = 10
DEFAULT_SEPARATION_RAYLEIGH = 3.0
DEFAULT_HARMONIC_MAX = 8
DEFAULT_DOWNSAMPLE_POINTS = 2000

#: Settings field that defaults each peak-selection parameter.
PEAK_DEFAULT_FIELDS: Mapping[str, str] = {
    "n": "n_peaks",
    "min_separation_rayleigh": "peak_separation_rayleigh",
    "alias_tolerance": "alias_freq_tolerance",
    "harmonic_max": "harmonic_max",
    "n_points": "downsample_points",
}


def peak_defaults_from_settings(settings: object | None) -> dict[str, Any]:
    """The peak-selection defaults a settings object carries."""
    out: dict[str, Any] = {}
    if settings is None:
        return out
    for parameter, field_name in PEAK_DEFAULT_FIELDS.items():
        value = getattr(settings, field_name, None)
        if value is not None:
            out[parameter] = value
    return out


@dataclass(frozen=True)
class Peak:
    """One periodogram peak; ``extra`` holds method scalars (BLS: depth, depth_snr,
    duration, t0, sde)."""

    period: float
    frequency: float
    power: float
    rank: int
    extra: Mapping[str, float] = field(default_factory=dict, hash=False)

    def to_dict(self) -> dict[str, float | int]:
        out: dict[str, float | int] = {
            "period": self.period,
            "frequency": self.frequency,
            "power": self.power,
            "rank": self.rank,
        }
        out.update(self.extra)
        return out


@dataclass(frozen=True, eq=False)
class Periodogram:
    """A computed periodogram of one light curve and one method, ascending in
    frequency (``period = 1/frequency``)."""

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
        if self.objective_sense not in ("max", "min"):
            raise ValueError(
                "objective_sense must be 'max' or 'min', got "
                f"{self.objective_sense!r}"
            )

    @classmethod
    def from_spectrum(
        cls,
        *,
        method: str,
        backend: str,
        frequency: np.ndarray,
        power: np.ndarray,
        objective_sense: ObjectiveSense,
        n_samples: int,
        baseline: float,
        extras: Mapping[str, np.ndarray] | None = None,
        meta: Mapping[str, Any] | None = None,
        peak_defaults: Mapping[str, Any] | None = None,
        order: np.ndarray | None = None,
    ) -> "Periodogram":
        """Build a periodogram, sorting everything into ascending frequency.

        ``order`` may pass the stable ascending-frequency permutation when the caller
        already knows it (a grid assembled in descending frequency is simply
        reversed); it must equal ``np.argsort(frequency, kind="stable")``.
        """
        if objective_sense not in ("max", "min"):
            raise ValueError(
                f"from_spectrum: objective_sense must be 'max' or 'min', got "
                f"{objective_sense!r}"
            )
        frequency = np.ascontiguousarray(frequency, dtype=np.float64)
        power = np.ascontiguousarray(power, dtype=np.float64)
        if frequency.ndim != 1:
            raise ValueError(
                f"from_spectrum: frequency must be 1-D, got shape {frequency.shape}"
            )
        if power.shape != frequency.shape:
            raise ValueError(
                f"from_spectrum: power shape {power.shape} does not match frequency "
                f"shape {frequency.shape}"
            )
        converted: dict[str, np.ndarray] = {}
        for key, values in (extras or {}).items():
            array = np.ascontiguousarray(values, dtype=np.float64)
            if array.shape != frequency.shape:
                raise ValueError(
                    f"from_spectrum: extras[{key!r}] shape {array.shape} does not "
                    f"match frequency shape {frequency.shape}"
                )
            converted[key] = array
        if order is None:
            order = np.argsort(frequency, kind="stable")
        frequency = frequency[order]
        power = power[order]
        sorted_extras = {k: v[order] for k, v in converted.items()}
        with np.errstate(divide="ignore"):
            period = 1.0 / frequency
        return cls(
            method=method,
            backend=backend,
            frequency=frequency,
            period=period,
            power=power,
            objective_sense=objective_sense,
            n_samples=n_samples,
            baseline=baseline,
            extras=sorted_extras,
            meta=dict(meta or {}),
            peak_defaults=dict(peak_defaults or {}),
        )

    def with_peak_defaults(self, defaults: Mapping[str, Any] | object) -> "Periodogram":
        """A copy whose peak-selection defaults are updated from a mapping or a
        settings object."""
        new = (
            dict(defaults)
            if isinstance(defaults, Mapping)
            else peak_defaults_from_settings(defaults)
        )
        unknown = sorted(set(new) - set(PEAK_DEFAULT_FIELDS))
        if unknown:
            raise ValueError(
                f"unknown peak default(s) {unknown}; expected a subset of "
                f"{sorted(PEAK_DEFAULT_FIELDS)}"
            )
        return dataclasses.replace(self, peak_defaults={**self.peak_defaults, **new})

    def _peak_default(self, name: str, fallback: Any) -> Any:
        value = self.peak_defaults.get(name)
        return fallback if value is None else value

    @property
    def size(self) -> int:
        return int(self.frequency.size)

    def _score(self) -> np.ndarray:
        return self.power if self.objective_sense == "max" else -self.power

    def best_periods(
        self,
        n: int | None = None,
        *,
        min_separation_rayleigh: float | None = None,
        alias_diverse: bool = False,
        harmonic_max: int | None = None,
        alias_tolerance: float | None = None,
    ) -> list[Peak]:
        """The ``n`` most significant peaks, most significant first."""
        count = int(self._peak_default("n", DEFAULT_N_PEAKS)) if n is None else int(n)
        if count <= 0 or self.size == 0:
            return []
        score = self._score()
        candidates = self._peak_candidates(score)
        if candidates.size == 0:
            return []
        chosen = self.select_peaks(
            self.frequency, score, candidates, count,
            min_separation_rayleigh=min_separation_rayleigh, alias_diverse=alias_diverse,
            harmonic_max=harmonic_max, alias_tolerance=alias_tolerance,
        )
        return [self._make_peak(int(idx), rank) for rank, idx in enumerate(chosen, 1)]

    def select_peaks(
        self,
        frequency: np.ndarray,
        score: np.ndarray,
        candidates: np.ndarray,
        count: int,
        *,
        min_separation_rayleigh: float | None = None,
        alias_diverse: bool = False,
        harmonic_max: int | None = None,
        alias_tolerance: float | None = None,
        examined: list | None = None,
    ) -> np.ndarray:
        """The greedy selection of :meth:`best_periods` on given candidates, with this
        periodogram's baseline and peak defaults (``examined`` receives the number of
        candidates visited)."""
        separation_rayleigh = float(
            self._peak_default("min_separation_rayleigh", DEFAULT_SEPARATION_RAYLEIGH)
            if min_separation_rayleigh is None
            else min_separation_rayleigh
        )
        resolved = self.baseline > 0.0
        separation = separation_rayleigh / self.baseline if resolved else 0.0
        if alias_diverse:
            order = int(
                self._peak_default("harmonic_max", DEFAULT_HARMONIC_MAX)
                if harmonic_max is None
                else harmonic_max
            )
            tolerance = (
                self._peak_default("alias_tolerance", None)
                if alias_tolerance is None
                else alias_tolerance
            )
            if tolerance is None:
                tolerance = min(separation, 1.0 / self.baseline) if resolved else 0.0
            return select_diverse_peaks(
                frequency, score, candidates, count,
                max(float(tolerance), 1e-12), order,
                min_separation=max(separation, 1e-12), examined=examined,
            )
        return select_top_peaks(frequency, score, candidates, count, separation,
                                examined=examined)

    def _peak_candidates(self, score: np.ndarray) -> np.ndarray:
        """Interior local maxima plus boundary samples exceeding their neighbour;
        non-finite scores and periods never qualify."""
        if _sum_is_finite(score) and self._periods_all_finite():
            safe = score
        else:
            usable = np.isfinite(score) & np.isfinite(self.period)
            safe = np.where(usable, score, -np.inf)
        candidates = _interior_maxima(safe)
        boundary: list[int] = []
        if self.size == 1:
            boundary.append(0)
        elif self.size >= 2:
            if safe[0] >= safe[1]:
                boundary.append(0)
            if safe[-1] > safe[-2]:
                boundary.append(self.size - 1)
        if boundary:
            candidates = np.concatenate(
                [candidates, np.asarray(boundary, dtype=np.int64)]
            )
        if candidates.size:
            ok = np.isfinite(score[candidates]) & np.isfinite(self.period[candidates])
            candidates = candidates[ok]
        if candidates.size == 0:
            finite = np.flatnonzero(np.isfinite(score) & np.isfinite(self.period))
            if finite.size == 0:
                return np.empty(0, dtype=np.int64)
            return finite[[int(np.argmax(score[finite]))]]
        return candidates

    def _periods_all_finite(self) -> bool:
        freq = self.frequency
        n = int(freq.size)
        if n == 0:
            return True
        if not (np.isfinite(freq[0]) and np.isfinite(freq[-1])):
            return False
        at_zero = int(np.searchsorted(freq, 0.0))
        if at_zero < n and freq[at_zero] == 0.0:
            return False
        return self.period.size == n

    def _make_peak(self, idx: int, rank: int) -> Peak:
        extra = {k: float(v[idx]) for k, v in self.extras.items()}
        return Peak(
            period=float(self.period[idx]),
            frequency=float(self.frequency[idx]),
            power=float(self.power[idx]),
            rank=rank,
            extra=extra,
        )

    def best_period(self) -> float:
        """The single most significant period in days (NaN if none is finite)."""
        peaks = self.best_periods(1)
        return peaks[0].period if peaks else float("nan")

    def downsample(self, n_points: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        """Peak-preserving downsample of ``(frequency, power)``."""
        count = (
            int(self._peak_default("n_points", DEFAULT_DOWNSAMPLE_POINTS))
            if n_points is None
            else int(n_points)
        )
        return peak_preserving_downsample(
            self.frequency, self.power, count, sense=self.objective_sense
        )

    def to_dict(
        self, *, n_best: int | None = None, include_raw: bool = False
    ) -> dict[str, Any]:
        out: dict[str, Any] = {
            "method": self.method,
            "backend": self.backend,
            "n_samples": self.n_samples,
            "baseline": self.baseline,
            "objective_sense": self.objective_sense,
            "peaks": [p.to_dict() for p in self.best_periods(n_best)],
        }
        if include_raw:
            freq, power = self.downsample()
            out["frequency"] = freq.tolist()
            out["power"] = power.tolist()
        return out


def _normalize_method_key(key: str) -> str:
    return "".join(ch for ch in key if ch.isalnum()).upper()


@dataclass(frozen=True, eq=False)
class MultiResult:
    """Several methods' periodograms of one light curve; lookups ignore case and
    separators."""

    results: Mapping[str, Periodogram]

    def _resolve(self, method: object) -> str | None:
        if not isinstance(method, str):
            return None
        if method in self.results:
            return method
        wanted = _normalize_method_key(method)
        for key in self.results:
            if _normalize_method_key(key) == wanted:
                return key
        return None

    def __getitem__(self, method: str) -> Periodogram:
        key = self._resolve(method)
        if key is None:
            raise KeyError(method)
        return self.results[key]

    def __contains__(self, method: object) -> bool:
        return self._resolve(method) is not None

    def __iter__(self) -> Iterator[str]:
        return iter(self.results)

    def __len__(self) -> int:
        return len(self.results)

    def get(self, method: str, default: Periodogram | None = None) -> Periodogram | None:
        key = self._resolve(method)
        return default if key is None else self.results[key]

    def keys(self):
        return self.results.keys()

    def values(self):
        return self.results.values()

    def items(self):
        return self.results.items()

    def best_periods(self, n: int | None = None, **kwargs: Any) -> dict[str, list[Peak]]:
        return {k: v.best_periods(n, **kwargs) for k, v in self.results.items()}

    def to_dict(self, *, n_best: int | None = None, include_raw: bool = False) -> dict:
        return {
            k: v.to_dict(n_best=n_best, include_raw=include_raw)
            for k, v in self.results.items()
        }


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
