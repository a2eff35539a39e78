"""Light curves: the input of every BLS search.

The subset of cuPeriod's ``LightCurve`` the BLS path reads
(src/cuperiod/core/lightcurve.py:58-303, 625-646): construction from arrays, the
``finite`` filter (non-finite samples and non-positive errors dropped), the
magnitude -> flux conversion BLS applies before searching (a transit is a dip in flux),
``baseline`` and the unusable-error diagnosis. Times are kept verbatim; the search
subtracts its own time origin in FP64.

:class:`MultiBandLightCurve` holds several bands of one star for the multi-band search
(src/cuperiod/core/lightcurve.py:329-575): every band a :class:`LightCurve`, all in one
brightness domain, in insertion order.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Mapping

import numpy as np

from cuperiod_amd.core.columns import Domain
from cuperiod_amd.core.errors import InsufficientDataError


def _as_f64(a: object, name: str) -> np.ndarray:
    if isinstance(a, np.ma.MaskedArray):
        a = a.filled(np.nan)
    out = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if out.ndim == 0:
        out = out.reshape(1)
    if out.ndim != 1:
        raise ValueError(f"{name} must be 1-D, got shape {out.shape}")
    return out


@dataclass(frozen=True, eq=False)
class LightCurve:
    """One band's ``time`` (days), ``value`` (magnitude or flux) and optional 1-sigma
    ``error``, as contiguous float64 arrays."""

    time: np.ndarray
    value: np.ndarray
    error: np.ndarray | None = None
    domain: Domain = Domain.MAGNITUDE
    meta: Mapping[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        time = _as_f64(self.time, "time")
        value = _as_f64(self.value, "value")
        if time.size != value.size:
            raise ValueError(
                f"time and value length mismatch: {time.size} != {value.size}"
            )
        error = None if self.error is None else _as_f64(self.error, "error")
        if error is not None and error.size != time.size:
            raise ValueError(
                f"error and time length mismatch: {error.size} != {time.size}"
            )
        object.__setattr__(self, "time", time)
        object.__setattr__(self, "value", value)
        object.__setattr__(self, "error", error)
        object.__setattr__(self, "domain", Domain(self.domain))

    @property
    def n(self) -> int:
        """Number of points."""
        return int(self.time.size)

    @property
    def baseline(self) -> float:
        """``max(time) - min(time)`` in days (0.0 if empty)."""
        if self.time.size == 0:
            return 0.0
        return float(self.time.max() - self.time.min())

    @classmethod
    def from_arrays(
        cls,
        time: object,
        value: object,
        error: object | None = None,
        *,
        domain: Domain | str = Domain.MAGNITUDE,
        meta: Mapping[str, Any] | None = None,
    ) -> "LightCurve":
        """Build from array-likes (masked entries become NaN)."""
        return cls(
            time=_as_f64(time, "time"),
            value=_as_f64(value, "value"),
            error=None if error is None else _as_f64(error, "error"),
            domain=Domain(domain),
            meta=dict(meta or {}),
        )

    def _usable(self) -> np.ndarray:
        mask = np.isfinite(self.time) & np.isfinite(self.value)
        if self.error is not None:
            mask &= np.isfinite(self.error) & (self.error > 0.0)
        return mask

    def finite(self) -> "LightCurve":
        """A copy without non-finite samples and without non-positive errors."""
        mask = self._usable()
        if mask.all():
            return self
        return LightCurve(
            time=self.time[mask],
            value=self.value[mask],
            error=None if self.error is None else self.error[mask],
            domain=self.domain,
            meta=self.meta,
        )

    def as_flux(self, *, zeropoint: float = 0.0) -> "LightCurve":
        """``flux = 10**(-0.4 (mag - zp))``, ``flux_err = 0.4 ln10 * flux * mag_err``."""
        if self.domain is Domain.FLUX:
            return self
        flux = np.power(10.0, -0.4 * (self.value - zeropoint))
        error = None if self.error is None else (0.4 * np.log(10.0)) * flux * self.error
        return LightCurve(self.time, flux, error, Domain.FLUX, self.meta)

    def as_magnitude(self, *, zeropoint: float = 0.0) -> "LightCurve":
        """``mag = -2.5 log10(flux) + zp``; non-positive flux becomes non-finite."""
        if self.domain is Domain.MAGNITUDE:
            return self
        with np.errstate(invalid="ignore", divide="ignore"):
            mag = -2.5 * np.log10(self.value) + zeropoint
            error = (
                None
                if self.error is None
                else (2.5 / np.log(10.0)) * self.error / self.value
            )
        return LightCurve(self.time, mag, error, Domain.MAGNITUDE, self.meta)

    def in_domain(self, domain: Domain | str, *, zeropoint: float = 0.0) -> "LightCurve":
        """This light curve converted to ``domain``."""
        if Domain(domain) is Domain.FLUX:
            return self.as_flux(zeropoint=zeropoint)
        return self.as_magnitude(zeropoint=zeropoint)

    def unusable_reason(self) -> str | None:
        """Why the errors leave no usable point, or ``None`` when some point survives."""
        if self.error is None:
            return None
        data = np.isfinite(self.time) & np.isfinite(self.value)
        n_data = int(np.count_nonzero(data))
        if n_data == 0 or self._usable().any():
            return None
        return (
            f"none of the {n_data} points with a finite time and value has a usable "
            "error (all are zero, negative or non-finite), so every point would be "
            "dropped; drop the error column (error=None) to run unweighted"
        )


@dataclass(frozen=True, eq=False)
class MultiBandLightCurve:
    """Several bands (filters) of one star: ``bands`` maps band name -> light curve, in
    insertion order; ``meta`` is carried through results."""

    bands: Mapping[str, LightCurve]
    meta: Mapping[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        if not self.bands:
            raise ValueError("MultiBandLightCurve requires at least one band")
        for name, lc in self.bands.items():
            if not isinstance(lc, LightCurve):
                raise TypeError(
                    f"band {name!r} is a {type(lc).__name__}, not a LightCurve; "
                    "build each band with LightCurve.from_arrays(...)"
                )
        domains = {name: lc.domain for name, lc in self.bands.items()}
        if len(set(domains.values())) > 1:
            listing = ", ".join(f"{k}: {v.value}" for k, v in domains.items())
            raise ValueError(
                f"bands are in different brightness domains ({listing}); convert them "
                "with LightCurve.in_domain(...), or pass domain='magnitude' or "
                "domain='flux' to periodogram()"
            )

    @property
    def band_names(self) -> tuple[str, ...]:
        """The band labels, in order."""
        return tuple(self.bands)

    @property
    def n_bands(self) -> int:
        """Number of bands."""
        return len(self.bands)

    @classmethod
    def from_light_curves(
        cls, bands: Mapping[str, LightCurve], *, meta: Mapping[str, Any] | None = None
    ) -> "MultiBandLightCurve":
        """Build from an explicit ``{band: LightCurve}`` mapping."""
        return cls(bands=dict(bands), meta=dict(meta or {}))

    def finite(self) -> "MultiBandLightCurve":
        """Every band without its non-finite samples and non-positive errors."""
        return MultiBandLightCurve(
            bands={k: lc.finite() for k, lc in self.bands.items()}, meta=self.meta
        )

    def in_domain(self, domain: Domain | str) -> "MultiBandLightCurve":
        """Every band converted to ``domain``."""
        return MultiBandLightCurve(
            bands={k: lc.in_domain(domain) for k, lc in self.bands.items()}, meta=self.meta
        )

    def unusable_reason(self) -> str | None:
        """Why the errors leave no usable point in any band, or ``None``: a band wiped
        out by its errors does not condemn the others."""
        reasons: dict[str, str] = {}
        for name, lc in self.bands.items():
            reason = lc.unusable_reason()
            if reason is not None:
                reasons[name] = reason
            elif lc._usable().any():
                return None
        if not reasons:
            return None
        return "; ".join(f"band {name!r}: {why}" for name, why in reasons.items())

    def stacked(self) -> tuple[np.ndarray, np.ndarray, np.ndarray | None, np.ndarray]:
        """All bands as time-sorted ``(time, value, error, band)`` arrays (stable sort;
        ``error`` is ``None`` when a band has none)."""
        times, values, errors, labels = [], [], [], []
        missing = False
        for name, lc in self.bands.items():
            times.append(lc.time)
            values.append(lc.value)
            if lc.error is None:
                missing = True
            else:
                errors.append(lc.error)
            labels.append(np.full(lc.n, name, dtype=object))
        time = np.concatenate(times)
        value = np.concatenate(values)
        band = np.concatenate(labels)
        error = None if missing else np.concatenate(errors)
        order = np.argsort(time, kind="stable")
        return (time[order], value[order], None if error is None else error[order],
                band[order])


def raise_if_unusable(data: "LightCurve | MultiBandLightCurve", label: str) -> None:
    """Raise :class:`InsufficientDataError` when an unusable error column would empty
    ``data`` (src/cuperiod/core/lightcurve.py:625-646)."""
    reason = data.unusable_reason()
    if reason is not None:
        raise InsufficientDataError(f"{label}: {reason}")


__all__ = ["LightCurve", "MultiBandLightCurve", "raise_if_unusable"]
