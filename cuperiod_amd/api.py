"""Single light-curve API: :func:`periodogram` (src/cuperiod/api.py:47-307).

Accepts a :class:`LightCurve`, a :class:`MultiBandLightCurve` or ``{band: LightCurve}``
mapping, or a ``(t, y[, dy])`` tuple (read as magnitudes unless ``domain`` says
otherwise, as in cuPeriod), one or several method names, and optional settings / grid /
backend; returns a :class:`Periodogram` (one method) or a :class:`MultiResult`.
"""

from __future__ import annotations

import warnings
from collections.abc import Mapping, Sequence
from typing import Any

import numpy as np

from cuperiod_amd.core.columns import Domain, coerce_domain
from cuperiod_amd.core.grid import GridSpec
from cuperiod_amd.core.lightcurve import LightCurve, MultiBandLightCurve, raise_if_unusable
from cuperiod_amd.core.result import MultiResult, Periodogram, peak_defaults_from_settings
from cuperiod_amd.methods.base import PeriodogramMethod, get_method, select_settings

_NORMALISED_FLUX_RANGE = (0.5, 1.5)


def _looks_like_arrays(obj: Any) -> bool:
    if not isinstance(obj, (tuple, list)) or not (2 <= len(obj) <= 3):
        return False
    first = obj[0]
    return hasattr(first, "__len__") and not isinstance(first, (str, bytes))


def _check_bands(bands: Mapping[str, Any], domain: Domain | None) -> None:
    """Every band a :class:`LightCurve`, all in one brightness domain unless ``domain``
    converts them (src/cuperiod/api.py:109-133)."""
    for name, lc in bands.items():
        if not isinstance(lc, LightCurve):
            raise TypeError(
                f"band {name!r} is a {type(lc).__name__}, not a LightCurve; build each "
                "band with LightCurve.from_arrays(...)"
            )
    if domain is None:
        domains = {name: lc.domain for name, lc in bands.items()}
        if len(set(domains.values())) > 1:
            listing = ", ".join(f"{k}: {v.value}" for k, v in domains.items())
            raise ValueError(
                f"bands are in different brightness domains ({listing}); pass "
                "domain='magnitude' or domain='flux' to convert them all, or convert the "
                "bands with LightCurve.in_domain(...)"
            )


def to_input(
    data: Any, *, domain: Domain | str | None = None
) -> LightCurve | MultiBandLightCurve:
    """A :class:`LightCurve` from a light curve or a ``(t, y[, dy])`` tuple, or a
    :class:`MultiBandLightCurve` from one or a ``{band: LightCurve}`` mapping."""
    dom = coerce_domain(domain)
    if isinstance(data, LightCurve):
        return data if dom is None else data.in_domain(dom)
    if isinstance(data, MultiBandLightCurve):
        _check_bands(data.bands, dom)
        return data if dom is None else data.in_domain(dom)
    if isinstance(data, Mapping):
        values = list(data.values())
        if values and all(isinstance(v, LightCurve) for v in values):
            _check_bands(data, dom)
            bands = data if dom is None else {k: lc.in_domain(dom) for k, lc in data.items()}
            return MultiBandLightCurve.from_light_curves(bands)
    if _looks_like_arrays(data):
        t, v = data[0], data[1]
        e = data[2] if len(data) == 3 else None
        return LightCurve.from_arrays(t, v, e, domain=dom or Domain.MAGNITUDE)
    raise TypeError(
        f"unsupported light-curve input {type(data).__name__}; pass a LightCurve, a "
        "MultiBandLightCurve, a {band: LightCurve} mapping or a (t, y[, dy]) tuple"
    )


def _flux_read_as_magnitudes(lc: LightCurve, methods: Sequence[PeriodogramMethod]) -> None:
    """A ``(t, y[, dy])`` tuple without ``domain`` is read as magnitudes. When a method
    that searches in flux is asked to, and the values sit around 1 the way a normalised
    flux does, say so: a dip in flux is a rise in magnitude, and the search would look
    for the wrong sign."""
    in_flux = [m.name for m in methods if m.natural_domain is Domain.FLUX]
    finite = lc.value[np.isfinite(lc.value)]
    if not in_flux or finite.size == 0:
        return
    centre = float(np.median(finite))
    if _NORMALISED_FLUX_RANGE[0] < centre < _NORMALISED_FLUX_RANGE[1]:
        warnings.warn(
            f"the values of the (t, y, dy) tuple have median {centre:.3g}, like a normalised "
            f"flux, yet are read as magnitudes for {'/'.join(in_flux)}; say which with "
            "domain='flux' or domain='magnitude'", UserWarning, stacklevel=3)


def _multiband_grid(method: PeriodogramMethod, mblc: MultiBandLightCurve, settings: Any
                    ) -> GridSpec:
    """A method's default grid from the time-sorted stack of every band
    (src/cuperiod/api.py:195-200)."""
    time, value, error, _ = mblc.finite().stacked()
    return method.default_grid(LightCurve.from_arrays(time, value, error), settings)


def _run_one(
    method: PeriodogramMethod,
    lc: LightCurve | MultiBandLightCurve,
    settings: Any,
    backend: str | None,
    grid: GridSpec | None,
) -> Periodogram:
    resolved = method.resolve_backend_for(backend, settings)
    if isinstance(lc, MultiBandLightCurve):
        if not method.supports_multiband:
            raise ValueError(f"{method.name} does not support multi-band input")
        raise_if_unusable(lc, f"{method.name} multiband")
        g = grid or _multiband_grid(method, lc, settings)
        pg = method.multiband_power(g, lc, settings, resolved)
    else:
        single = lc.in_domain(method.natural_domain) if method.natural_domain else lc
        raise_if_unusable(single, method.name)
        g = grid or method.default_grid(single, settings)
        pg = method.power(g, single, settings, resolved, engine=None)
    return pg.with_peak_defaults(peak_defaults_from_settings(settings))


def periodogram(
    data: Any,
    method: str | Sequence[str] = "BLS",
    *,
    backend: str | None = "auto",
    grid: GridSpec | None = None,
    settings: Any = None,
    domain: Domain | str | None = None,
) -> Periodogram | MultiResult:
    """Periodogram(s) of one light curve: one method name gives a :class:`Periodogram`,
    a list of names a :class:`MultiResult` keyed by method. ``backend="auto"`` defers to
    the settings' ``backend`` and resolves to ``"hip"`` on an AMD GPU."""
    lc = to_input(data, domain=domain)
    single = isinstance(method, str)
    methods = [get_method(name) for name in ([method] if single else method)]
    per_method = select_settings(methods, settings)
    if domain is None and _looks_like_arrays(data):
        _flux_read_as_magnitudes(lc, methods)
    found = {m.name: _run_one(m, lc, per_method[m.name], backend, grid) for m in methods}
    return found[methods[0].name] if single else MultiResult(found)


def best_periods(result: Periodogram | MultiResult, n: int | None = None, **kwargs: Any):
    """``result.best_periods(n)``."""
    return result.best_periods(n, **kwargs)


__all__ = ["best_periods", "periodogram", "to_input"]
