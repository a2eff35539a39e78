"""Multi-band Box Least Squares on the MI355X engine.

Bands of one star share one trial *period*. Two statistics
(``BLSSettings.mb_ephemeris`` or the ``ephemeris=`` argument of
:func:`bls_multiband_power`; cuPeriod src/cuperiod/multiband/bls_mb.py:1-35):

``"independent"`` (default)
    Each band is box-searched on its own over the common grid (its own best duration
    and mid-transit time per period); the per-period statistic is the quadrature sum of
    the bands' best depth-SNRs, ``sum_k max(SNR_k, 0)^2``. The box geometry at each
    period is that of the band detecting it most strongly.

``"shared"``
    One box -- one duration and one mid-transit time on a phase grid common to every
    band -- is placed for all bands at once, each band fitting its own depth inside it;
    the summed squared depth-SNRs of the bands that dip there are maximised over the
    boxes (:mod:`cuperiod_amd.methods._bls_joint`). Duration and mid-transit time are
    the shared box's.

Either way the power is ``sum_k max(SNR_k, 0)^2``, ``extras["band_index"]`` indexes
``meta["bands"]`` (the contributing bands), the depth is that band's in its own flux
units, and ``meta["mb_ephemeris"]`` names the statistic. A band's errors are used only
when every contributing band has them; otherwise every band is searched with a
constant error equal to its own robust scatter (1.4826 x the median absolute deviation).
"""

from __future__ import annotations

from typing import Any, Final

import numpy as np

from cuperiod_amd.core.columns import Domain
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.core.errors import InsufficientDataError
from cuperiod_amd.core.grid import GridSpec
from cuperiod_amd.core.lightcurve import LightCurve, MultiBandLightCurve
from cuperiod_amd.core.result import Periodogram

#: A band needs at least this many finite points to contribute a box search.
_MIN_BAND_POINTS = 3

#: Scale factor turning a median absolute deviation into a Gaussian sigma.
_MAD_TO_SIGMA = 1.4826

#: The multi-band box statistics.
EPHEMERIS_MODES: Final = ("independent", "shared")


def _robust_scatter(flux: np.ndarray) -> float:
    """``1.4826 * MAD`` of a band's flux (bls_mb.py:63-76); the standard deviation when
    the MAD is 0, and 1 for a constant band."""
    sigma = _MAD_TO_SIGMA * float(np.median(np.abs(flux - np.median(flux))))
    if not (np.isfinite(sigma) and sigma > 0.0):
        sigma = float(np.std(flux))
    if not (np.isfinite(sigma) and sigma > 0.0):
        sigma = 1.0
    return sigma


def _band_errors(bands: list[LightCurve]) -> list[np.ndarray]:
    """Real errors when every band has them, else each band's constant robust scatter
    (bls_mb.py:79-92)."""
    if all(lc.error is not None for lc in bands):
        return [np.asarray(lc.error, dtype=np.float64) for lc in bands]
    return [np.full(lc.n, _robust_scatter(lc.value)) for lc in bands]


def resolve_ephemeris(settings: Any, ephemeris: str | None = None) -> str:
    """The statistic to run: the argument, else ``settings.mb_ephemeris``, else
    ``"independent"`` (bls_mb.py:95-117). Raises ``ValueError`` for another name."""
    if ephemeris is not None:
        mode = ephemeris
    elif isinstance(settings, BLSSettings):
        mode = settings.mb_ephemeris
    else:
        mode = "independent"
    mode = str(mode).strip().lower()
    if mode not in EPHEMERIS_MODES:
        raise ValueError(
            f"BLS multiband: unknown ephemeris {mode!r}; choose from {EPHEMERIS_MODES}"
        )
    return mode


def bls_multiband_power(
    grid: GridSpec,
    mblc: MultiBandLightCurve,
    settings: BLSSettings,
    backend: str,
    method: Any,
    *,
    ephemeris: str | None = None,
) -> Periodogram:
    """The multi-band BLS statistic on a shared period grid (bls_mb.py:120-237).

    ``grid`` is searched exactly as given, segmented as the single-band search segments
    it; a band contributes when it has at least 3 finite points. Raises
    :class:`InsufficientDataError` when no band contributes, the contributing bands hold
    fewer than ``min_detections`` points, or the grid or the baseline is empty, and
    ``ValueError`` for an unknown ``ephemeris``.
    """
    from cuperiod_amd.methods.bls import _grid_segments, _segments_power, sde_of

    mode = resolve_ephemeris(settings, ephemeris)
    finite = mblc.finite()
    used = [
        (name, lc.in_domain(Domain.FLUX))
        for name, lc in finite.bands.items()
        if lc.n >= _MIN_BAND_POINTS
    ]
    if not used:
        raise InsufficientDataError("BLS multiband: no band has enough finite points")
    bands = [lc for _, lc in used]
    n_total = sum(lc.n for lc in bands)
    if n_total < settings.min_detections:
        raise InsufficientDataError(
            f"BLS multiband: {n_total} finite points across usable bands "
            f"< min_detections {settings.min_detections}"
        )
    stacked_time = np.concatenate([lc.time for lc in bands])
    baseline = float(stacked_time.max() - stacked_time.min())
    if baseline <= 0.0:
        raise InsufficientDataError("BLS multiband: no usable time baseline")
    segments = _grid_segments(grid, settings)
    if not segments:
        raise InsufficientDataError("BLS multiband: empty trial grid")
    errors = _band_errors(bands)
    meta: dict[str, Any] = {
        **dict(finite.meta),
        "bands": tuple(name for name, _ in used),
        "mb_ephemeris": mode,
    }
    if mode == "shared":
        period, power, extras = _shared(bands, errors, segments, settings, backend)
    else:
        period, power, extras = _independent(bands, errors, segments, settings, backend,
                                             _segments_power)
    power = np.nan_to_num(power, nan=0.0, posinf=0.0)
    return Periodogram.from_spectrum(
        method="BLS",
        backend=backend,
        frequency=1.0 / period,
        power=power,
        objective_sense="max",
        n_samples=n_total,
        baseline=baseline,
        extras={
            "depth": extras["depth"],
            "depth_snr": extras["depth_snr"],
            "duration": extras["duration"],
            "t0": extras["t0"],
            "sde": sde_of(power),
            "band_index": extras["band_index"],
        },
        meta=meta,
    )


def _independent(bands, errors, segments, settings, backend, segments_power):
    """The independent per-band stack (bls_mb.py:240-275): ``(period, power, extras)``."""
    per_band = [
        segments_power(backend, lc.time, lc.value, err, segments, settings)
        for lc, err in zip(bands, errors, strict=True)
    ]
    parts: dict[str, list[np.ndarray]] = {
        name: [] for name in ("period", "power", "depth", "depth_snr", "duration", "t0",
                              "band_index")
    }
    for s, (periods, _) in enumerate(segments):
        seg = [band_segments[s] for band_segments in per_band]
        dsnr = np.vstack([np.clip(b["depth_snr"], 0.0, None) for b in seg])
        depth = np.vstack([b["depth"] for b in seg])
        dur = np.vstack([b["duration"] for b in seg])
        t0 = np.vstack([b["transit_time"] for b in seg])
        combined = np.sum(dsnr**2, axis=0)
        best = np.argmax(dsnr, axis=0)
        cols = np.arange(dsnr.shape[1])
        parts["period"].append(np.asarray(periods, dtype=np.float64))
        parts["power"].append(combined)
        parts["depth_snr"].append(np.sqrt(combined))
        parts["depth"].append(depth[best, cols])
        parts["duration"].append(dur[best, cols])
        parts["t0"].append(t0[best, cols])
        parts["band_index"].append(best.astype(np.float64))
    joined = {name: np.concatenate(values) for name, values in parts.items()}
    period, power = joined.pop("period"), joined.pop("power")
    return period, power, joined


def _shared(bands, errors, segments, settings, backend):
    """The joint-ephemeris search (bls_mb.py:278-312): ``(period, power, extras)``."""
    from cuperiod_amd.methods._bls_joint import joint_bands, joint_search

    prepared = joint_bands([lc.time for lc in bands], [lc.value for lc in bands], errors)
    found = joint_search(
        prepared,
        segments,
        settings.bins_per_duration,
        objective=settings.objective,
        backend=backend,
        batch=settings.batch_periods,
        precision=settings.precision,
    )
    period = np.concatenate([np.asarray(p, dtype=np.float64) for p, _ in segments])

    def joined(name: str) -> np.ndarray:
        return np.concatenate([f[name] for f in found])

    extras = {
        "depth": joined("depth"),
        "depth_snr": joined("depth_snr"),
        "duration": joined("duration"),
        "t0": joined("transit_time"),
        "band_index": joined("band_index"),
    }
    return period, joined("power"), extras


__all__ = ["EPHEMERIS_MODES", "bls_multiband_power", "resolve_ephemeris"]
