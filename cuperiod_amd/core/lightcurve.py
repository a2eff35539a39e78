"""Light curves: the inputs of the searches.

:class:`LightCurve` carries the fields and the handful of operations the BLS path reads
from cuPeriod's light curve (src/cuperiod/core/lightcurve.py:58-303): ``time`` (days),
``value`` (magnitude or flux), an optional 1-sigma ``error``, the brightness ``domain``
and ``meta``; :meth:`LightCurve.finite` keeps the points a search can use (finite time and
value, and a finite positive error when there is an error column); the domain
conversions are the photometric ones (``flux = 10**(-0.4 (mag - zp))``). Table and file
readers are not part of this package.

:class:`MultiBandLightCurve` groups several bands of one star (one brightness domain) for
the multi-band search (src/cuperiod/core/lightcurve.py:329-575).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Mapping

import numpy as np

from cuperiod_amd.core.columns import Domain
from cuperiod_amd.core.errors import InsufficientDataError

_LN10 = math.log(10.0)


def _column(data: object, label: str, length: int | None = None) -> np.ndarray:
    """``data`` as a contiguous 1-D float64 array (masked entries as NaN, a scalar as one
    value), of ``length`` values when given."""
    if isinstance(data, np.ma.MaskedArray):
        data = data.astype(np.float64).filled(np.nan)
    arr = np.atleast_1d(np.asarray(data, dtype=np.float64))
    if arr.ndim != 1:
        raise ValueError(f"light curve {label}: expected one dimension, got shape {arr.shape}")
    if length is not None and arr.size != length:
        raise ValueError(f"light curve {label}: {arr.size} values for {length} times")
    return np.ascontiguousarray(arr)


@dataclass(frozen=True, eq=False)
class LightCurve:
    """One band: ``time`` (days), ``value`` and optional ``error`` (float64 arrays of one
    length), its brightness ``domain`` and free-form ``meta``."""

    time: np.ndarray
    value: np.ndarray
    error: np.ndarray | None = None
    domain: Domain = Domain.MAGNITUDE
    meta: Mapping[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        time = _column(self.time, "time")
        set_ = object.__setattr__
        set_(self, "time", time)
        set_(self, "value", _column(self.value, "value", time.size))
        if self.error is not None:
            set_(self, "error", _column(self.error, "error", time.size))
        set_(self, "domain", Domain(self.domain))

    @classmethod
    def from_arrays(cls, time: object, value: object, error: object | None = None, *,
                    domain: Domain | str = Domain.MAGNITUDE,
                    meta: Mapping[str, Any] | None = None) -> "LightCurve":
        """A light curve from array-likes (masked entries become NaN, dropped by
        :meth:`finite`)."""
        return cls(time, value, error, Domain(domain), dict(meta or {}))

    @property
    def n(self) -> int:
        """Number of points."""
        return int(self.time.size)

    @property
    def baseline(self) -> float:
        """Span of the times in days (0.0 without points)."""
        return float(np.ptp(self.time)) if self.time.size else 0.0

    def _kept(self) -> np.ndarray:
        """The points a search keeps."""
        keep = np.isfinite(self.time) & np.isfinite(self.value)
        if self.error is not None:
            keep &= np.isfinite(self.error) & (self.error > 0.0)
        return keep

    def _subset(self, keep: np.ndarray) -> "LightCurve":
        return LightCurve(self.time[keep], self.value[keep],
                          None if self.error is None else self.error[keep], self.domain,
                          self.meta)

    def finite(self) -> "LightCurve":
        """The usable points only (this light curve when all are usable)."""
        keep = self._kept()
        return self if keep.all() else self._subset(keep)

    def in_domain(self, domain: Domain | str, *, zeropoint: float = 0.0) -> "LightCurve":
        """The light curve in ``domain``: ``flux = 10**(-0.4 (mag - zp))`` with
        ``sigma_flux = 0.4 ln10 flux sigma_mag``, and the inverse (non-positive flux has
        no magnitude: NaN)."""
        target = Domain(domain)
        if target is self.domain:
            return self
        if target is Domain.FLUX:
            value = 10.0 ** (-0.4 * (self.value - zeropoint))
            error = None if self.error is None else (0.4 * _LN10) * value * self.error
        else:
            with np.errstate(invalid="ignore", divide="ignore"):
                value = zeropoint - 2.5 * np.log10(self.value)
                error = None if self.error is None else (2.5 / _LN10) * self.error / self.value
        return LightCurve(self.time, value, error, target, self.meta)

    def as_flux(self, *, zeropoint: float = 0.0) -> "LightCurve":
        return self.in_domain(Domain.FLUX, zeropoint=zeropoint)

    def as_magnitude(self, *, zeropoint: float = 0.0) -> "LightCurve":
        return self.in_domain(Domain.MAGNITUDE, zeropoint=zeropoint)

    def unusable_reason(self) -> str | None:
        """Why the error column leaves no point to search, or None."""
        if self.error is None:
            return None
        with_data = int(np.count_nonzero(np.isfinite(self.time) & np.isfinite(self.value)))
        if with_data == 0 or self._kept().any():
            return None
        return (f"every one of the {with_data} measured points has an error that is zero, "
                "negative or not finite; pass error=None to search without weights")


@dataclass(frozen=True, eq=False)
class MultiBandLightCurve:
    """Bands (filters) of one star: ``bands`` maps a band name to its light curve, in the
    order given; every band is in one brightness domain."""

    bands: Mapping[str, LightCurve]
    meta: Mapping[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        if len(self.bands) == 0:
            raise ValueError("a multi-band light curve needs at least one band")
        wrong = [k for k, v in self.bands.items() if not isinstance(v, LightCurve)]
        if wrong:
            raise TypeError(f"bands {wrong} are not LightCurve objects "
                            "(make them with LightCurve.from_arrays)")
        seen = {lc.domain for lc in self.bands.values()}
        if len(seen) > 1:
            raise ValueError("the bands mix brightness domains "
                             f"({sorted(d.value for d in seen)}); convert them to one")

    @classmethod
    def from_light_curves(cls, bands: Mapping[str, LightCurve], *,
                          meta: Mapping[str, Any] | None = None) -> "MultiBandLightCurve":
        return cls(dict(bands), dict(meta or {}))

    @property
    def band_names(self) -> tuple[str, ...]:
        return tuple(self.bands.keys())

    @property
    def n_bands(self) -> int:
        return len(self.bands)

    def _each(self, op) -> "MultiBandLightCurve":
        return MultiBandLightCurve({k: op(v) for k, v in self.bands.items()}, self.meta)

    def finite(self) -> "MultiBandLightCurve":
        """Every band's usable points."""
        return self._each(LightCurve.finite)

    def in_domain(self, domain: Domain | str) -> "MultiBandLightCurve":
        return self._each(lambda lc: lc.in_domain(domain))

    def unusable_reason(self) -> str | None:
        """Why no band has a point to search, or None (one usable band suffices)."""
        if any(lc._kept().any() for lc in self.bands.values()):
            return None
        parts = [f"band {k!r}: {why}" for k, lc in self.bands.items()
                 if (why := lc.unusable_reason()) is not None]
        return "; ".join(parts) or None

    def stacked(self) -> tuple[np.ndarray, np.ndarray, np.ndarray | None, np.ndarray]:
        """``(time, value, error, band)`` of all bands merged in time order (stable on
        equal times; ``error`` None unless every band has one)."""
        lcs = list(self.bands.values())
        time = np.concatenate([lc.time for lc in lcs])
        order = np.argsort(time, kind="stable")
        value = np.concatenate([lc.value for lc in lcs])[order]
        band = np.concatenate([np.full(lc.n, k, dtype=object)
                               for k, lc in self.bands.items()])[order]
        error = None
        if all(lc.error is not None for lc in lcs):
            error = np.concatenate([lc.error for lc in lcs])[order]
        return time[order], value, error, band


def raise_if_unusable(data: LightCurve | MultiBandLightCurve, label: str) -> None:
    """:class:`InsufficientDataError` when the error column leaves nothing to search
    (the check of src/cuperiod/core/lightcurve.py:625-646)."""
    why = data.unusable_reason()
    if why is not None:
        raise InsufficientDataError(f"{label}: {why}")


__all__ = ["LightCurve", "MultiBandLightCurve", "raise_if_unusable"]
