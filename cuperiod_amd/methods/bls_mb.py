"""Multi-band Box Least Squares on the MI355X engine.

The bands of one star share a trial *period*; the statistic at each period is
``sum_k max(SNR_k, 0)^2`` over the bands' depth signal-to-noise ratios. Two ways to place
the boxes (``BLSSettings.mb_ephemeris`` or the ``ephemeris=`` argument; the modes of
cuPeriod's src/cuperiod/multiband/bls_mb.py:1-35):

``"independent"`` (default)
    every band runs its own single-band search on the common grid, and each period's
    statistic stacks the bands' best boxes; duration, mid-transit time and depth are
    those of the band with the largest SNR there (the first such band on equal SNRs).

``"shared"``
    one box (a duration and a mid-transit time on a phase grid common to all bands) is
    placed for every band at once, each band fitting its own depth in it, and the box
    with the largest stacked statistic wins (:mod:`cuperiod_amd.methods._bls_joint`,
    which drives the joint kernel).

``extras["band_index"]`` indexes ``meta["bands"]``, the bands that took part (those with
at least three usable points); ``meta["mb_ephemeris"]`` names the mode. Every band is
searched in flux. A band's errors are used when every taking-part band has errors;
otherwise each band gets one constant error, its robust scatter.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Final

import numpy as np

from cuperiod_amd.core.columns import Domain
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.core.errors import InsufficientDataError
from cuperiod_amd.core.grid import GridSpec
from cuperiod_amd.core.lightcurve import LightCurve, MultiBandLightCurve
from cuperiod_amd.core.result import Periodogram

#: The multi-band box statistics.
EPHEMERIS_MODES: Final = ("independent", "shared")

#: Usable points a band needs to take part.
_BAND_MIN_POINTS = 3

#: sigma / MAD of a normal distribution.
_SIGMA_PER_MAD = 1.4826

#: Per-period outputs of a band search, by the name the periodogram gives them.
_BOX_FIELDS = (("depth", "depth"), ("duration", "duration"), ("t0", "transit_time"))


def _noise_scale(flux: np.ndarray) -> float:
    """A band's constant error: ``1.4826 x MAD``; the standard deviation where the MAD
    is zero (more than half the points equal), and 1.0 for a constant band (the rule of
    src/cuperiod/multiband/bls_mb.py:63-76)."""
    mad = np.median(np.abs(flux - np.median(flux)))
    for scale in (_SIGMA_PER_MAD * float(mad), float(np.std(flux))):
        if np.isfinite(scale) and scale > 0.0:
            return scale
    return 1.0


def resolve_ephemeris(settings: Any, ephemeris: str | None = None) -> str:
    """The mode to run: ``ephemeris`` if given, else ``settings.mb_ephemeris`` for BLS
    settings, else ``"independent"``; names are trimmed and case-folded, and a name not in
    :data:`EPHEMERIS_MODES` raises ``ValueError``."""
    if ephemeris is None:
        ephemeris = settings.mb_ephemeris if isinstance(settings, BLSSettings) else None
    mode = "independent" if ephemeris is None else str(ephemeris).strip().lower()
    if mode not in EPHEMERIS_MODES:
        raise ValueError(f"BLS multiband: ephemeris {ephemeris!r} is none of {EPHEMERIS_MODES}")
    return mode


@dataclass(frozen=True)
class _BandSet:
    """The bands that take part: names, flux light curves and the error array each is
    searched with."""

    names: tuple[str, ...]
    curves: tuple[LightCurve, ...]
    errors: tuple[np.ndarray, ...]

    @classmethod
    def select(cls, mblc: MultiBandLightCurve, min_detections: int) -> "_BandSet":
        usable = [(name, lc.in_domain(Domain.FLUX)) for name, lc in mblc.finite().bands.items()]
        usable = [(name, lc) for name, lc in usable if lc.n >= _BAND_MIN_POINTS]
        if not usable:
            raise InsufficientDataError(
                f"BLS multiband: not one band has enough finite points (at least "
                f"{_BAND_MIN_POINTS} are needed)")
        names, curves = (tuple(col) for col in zip(*usable))
        total = sum(lc.n for lc in curves)
        if total < min_detections:
            raise InsufficientDataError(
                f"BLS multiband: the usable bands hold {total} finite points, fewer than "
                f"min_detections = {min_detections}")
        if all(lc.error is not None for lc in curves):
            errors = tuple(np.asarray(lc.error, dtype=np.float64) for lc in curves)
        else:
            errors = tuple(np.full(lc.n, _noise_scale(lc.value)) for lc in curves)
        return cls(names, curves, errors)

    @property
    def n_points(self) -> int:
        return sum(lc.n for lc in self.curves)

    @property
    def baseline(self) -> float:
        lo = min(float(lc.time.min()) for lc in self.curves)
        hi = max(float(lc.time.max()) for lc in self.curves)
        return hi - lo


def _stack_independent(per_band: list[list[dict]]) -> tuple[np.ndarray, dict]:
    """``(power, extras)`` over every period of the grid from each band's per-segment
    outputs (``per_band[band][segment]``)."""
    def matrix(key: str) -> np.ndarray:
        return np.stack([np.concatenate([seg[key] for seg in segs]) for segs in per_band])

    snr = np.clip(matrix("depth_snr"), 0.0, None)
    power = np.square(snr).sum(axis=0)
    lead = np.argmax(snr, axis=0)
    at_lead = lambda m: np.take_along_axis(m, lead[None, :], axis=0)[0]  # noqa: E731
    extras = {name: at_lead(matrix(key)) for name, key in _BOX_FIELDS}
    extras["depth_snr"] = np.sqrt(power)
    extras["band_index"] = lead.astype(np.float64)
    return power, extras


def _independent(bands: _BandSet, segments, settings: BLSSettings, backend: str):
    from cuperiod_amd.methods import bls

    per_band = [bls._segments_power(backend, lc.time, lc.value, err, segments, settings)
                for lc, err in zip(bands.curves, bands.errors)]
    return _stack_independent(per_band)


def _shared(bands: _BandSet, segments, settings: BLSSettings, backend: str):
    from cuperiod_amd.methods import _bls_joint

    joint = _bls_joint.joint_bands([lc.time for lc in bands.curves],
                                   [lc.value for lc in bands.curves], list(bands.errors))
    found = _bls_joint.joint_search(joint, segments, settings.bins_per_duration,
                                    objective=settings.objective, backend=backend,
                                    batch=settings.batch_periods,
                                    precision=settings.precision)
    column = lambda key: np.concatenate([f[key] for f in found])  # noqa: E731
    extras = {name: column(key) for name, key in _BOX_FIELDS}
    extras["depth_snr"] = column("depth_snr")
    extras["band_index"] = column("band_index")
    return column("power"), extras


def bls_multiband_power(
    grid: GridSpec,
    mblc: MultiBandLightCurve,
    settings: BLSSettings,
    backend: str,
    method: Any,
    *,
    ephemeris: str | None = None,
) -> Periodogram:
    """The multi-band statistic over ``grid``, segmented as the single-band search
    segments it (src/cuperiod/multiband/bls_mb.py:120-237). Raises
    :class:`InsufficientDataError` when no band has three usable points, the bands that
    do hold fewer than ``settings.min_detections``, the times span no baseline or the grid
    is empty, and ``ValueError`` for an unknown ``ephemeris``."""
    from cuperiod_amd.methods.bls import _grid_segments, sde_of

    mode = resolve_ephemeris(settings, ephemeris)
    bands = _BandSet.select(mblc, settings.min_detections)
    if not bands.baseline > 0.0:
        raise InsufficientDataError("BLS multiband: every usable point has the same time")
    segments = _grid_segments(grid, settings)
    if not segments:
        raise InsufficientDataError("BLS multiband: empty trial grid")
    run = _shared if mode == "shared" else _independent
    power, found = run(bands, segments, settings, backend)
    power = np.nan_to_num(power, nan=0.0, posinf=0.0)
    periods = np.concatenate([np.asarray(p, dtype=np.float64) for p, _ in segments])
    extras = {name: found[name] for name in ("depth", "depth_snr", "duration", "t0")}
    extras["sde"] = sde_of(power)
    extras["band_index"] = found["band_index"]
    return Periodogram.from_spectrum(
        method="BLS", backend=backend, frequency=1.0 / periods, power=power,
        objective_sense="max", n_samples=bands.n_points, baseline=bands.baseline,
        extras=extras,
        meta={**dict(mblc.meta), "bands": bands.names, "mb_ephemeris": mode})


__all__ = ["EPHEMERIS_MODES", "bls_multiband_power", "resolve_ephemeris"]
