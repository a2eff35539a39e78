"""Transit Least Squares (limb-darkened matched filter) on the MI355X engine.

cuPeriod's TLS (src/cuperiod/methods/tls.py:1-26): instead of BLS's rectangular box,
the folded light curve is correlated with a quadratic limb-darkened transit template
(Hippke & Heller 2019 in spirit) resampled to each trial duration; at every phase
position the out-of-transit level and the depth are fitted jointly by weighted least
squares. With ``y'`` the flux minus its weighted mean, ``g`` the template and
``W = sum w``, the depth is ``-(sum w g y')/D`` and the signal residue
``SR = (sum w g y')^2 / D`` with ``D = sum w g^2 - (sum w g)^2/W``. The per-period
maximum SR is normalised over the grid into the SDE (the periodogram's power). Each
period also reports the best template's ``duration`` and ``t0``, the mid-transit time
of the first transit at or after the first observation (the first observation where
no dip qualifies).

The search runs in one HIP launch per call (``cpb_tls_search``: exact 64-bit
fixed-point phase bins, FP64 correlation in the reference's order); the host keeps the
reference's preparation (flux scale, weighted mean, templates) and finish (SDE, units,
epoch).
"""

from __future__ import annotations

import ctypes
import math
from typing import ClassVar, Final

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.core.columns import Domain
from cuperiod_amd.core.config import TLSSettings
from cuperiod_amd.core.errors import BackendUnavailableError, InsufficientDataError
from cuperiod_amd.core.grid import GridSpec, check_grid_size
from cuperiod_amd.core.lightcurve import LightCurve
from cuperiod_amd.core.result import Periodogram
from cuperiod_amd.methods._bls_core import HIP_BACKEND, _device_of, _ptr, _stream_handle, _torch
from cuperiod_amd.methods.base import PeriodogramMethod, register, searchable_points

#: Smallest admissible template energy (an empty window is skipped).
_W_EPS: Final = 1e-300
#: Relative floor on the level-corrected template energy ``D`` against ``sum w g^2``.
_DEN_REL_EPS: Final = 1e-9
#: A flux within this many ulps of constant is constant.
_CONSTANT_ULPS: Final = 64.0
#: The per-period outputs.
_OUTPUTS: Final = ("sde", "sr", "depth", "duration", "t0")


def limb_darkened_template(n: int, u1: float, u2: float) -> np.ndarray:
    """Peak-normalised central-transit deficit profile over ``n`` samples: the quadratic
    law ``1 - u1(1-mu) - u2(1-mu)^2`` at the bin centres of the chord
    (src/cuperiod/methods/tls.py:81-109)."""
    if n <= 1:
        return np.ones(max(n, 1), dtype=np.float64)
    x = (np.arange(n, dtype=np.float64) + 0.5) / n - 0.5
    z = np.clip(2.0 * np.abs(x), 0.0, 1.0)
    mu = np.sqrt(1.0 - z * z)
    g = 1.0 - u1 * (1.0 - mu) - u2 * (1.0 - mu) ** 2
    g = np.where(z < 1.0, g, 0.0)
    peak = float(g.max())
    return g / peak if peak > 0.0 else g


def _duration_bins(settings: TLSSettings) -> list[int]:
    """Template widths in phase bins, deduplicated, ascending, clamped to
    ``[1, n_bins-1]`` (src/cuperiod/methods/tls.py:112-126)."""
    n_bins = settings.n_phase_bins
    fracs = np.geomspace(settings.duration_min_frac, settings.duration_max_frac,
                         settings.n_durations)
    return sorted({min(max(int(round(f * n_bins)), 1), n_bins - 1) for f in fracs})


def _grid_duration_frac(settings: TLSSettings) -> float:
    """Phase drift per grid step: ``grid_duration_frac``, at most twice the shortest
    duration (src/cuperiod/methods/tls.py:129-140)."""
    return min(settings.grid_duration_frac, 2.0 * settings.duration_min_frac)


def _period_grid(baseline: float, settings: TLSSettings, *, time: np.ndarray | None = None
                 ) -> np.ndarray:
    """Ascending trial periods of ``np.arange(1/max_period, 1/min_period, df)``, the size
    checked before allocation (src/cuperiod/methods/tls.py:143-176)."""
    max_period = min(settings.max_period_days, baseline / settings.min_transits)
    if max_period <= settings.min_period_days:
        return np.zeros(0, dtype=np.float64)
    df = _grid_duration_frac(settings) / (settings.oversample * baseline)
    f_lo, f_hi = 1.0 / max_period, 1.0 / settings.min_period_days
    check_grid_size(
        math.ceil((f_hi - f_lo) / df), automatic=True, time=time, baseline=baseline,
        maximum_frequency=f_hi,
        context=("TLS default period grid (min_period_days, grid_duration_frac and "
                 "oversample set its size)"),
    )
    freq = np.arange(f_lo, f_hi, df)
    if freq.size == 0:
        return np.zeros(0, dtype=np.float64)
    return np.ascontiguousarray(1.0 / freq[::-1], dtype=np.float64)


def _is_constant(y: np.ndarray) -> bool:
    """Whether the flux is constant to within 64 ulps of its largest magnitude
    (src/cuperiod/methods/tls.py:185-197)."""
    if y.size == 0:
        return True
    scale = float(np.max(np.abs(y)))
    return float(np.ptp(y)) <= _CONSTANT_ULPS * float(np.finfo(np.float64).eps) * scale


def _flux_scale(y: np.ndarray, dy: np.ndarray | None) -> float:
    """A power of two near the flux noise scale (``median(dy)``, else ``std(y)``)
    (src/cuperiod/methods/tls.py:200-213)."""
    s = float(np.std(y)) if dy is None else float(np.median(np.abs(dy)))
    if not (math.isfinite(s) and s > 0.0):
        return 1.0
    return 2.0 ** min(max(round(math.log2(s)), -1000), 1000)


def _no_transit(n_periods: int, t_ref: float) -> dict[str, np.ndarray]:
    out = {key: np.zeros(n_periods, dtype=np.float64) for key in _OUTPUTS}
    out["t0"] = np.full(n_periods, t_ref, dtype=np.float64)
    return out


def _tls_hip(tau, yw, w, periods, *, n_bins, dur_bins, templates, w_total, den_rel):
    """The matched filter on the device: ``{sr, depth, duration, t0}`` (t0 from
    ``tau = 0``)."""
    return _tls_hip_curves([(tau, yw, w, w_total, periods)], n_bins=n_bins,
                           dur_bins=dur_bins, templates=templates, den_rel=den_rel)[0]


def _tls_hip_curves(curves, *, n_bins, dur_bins, templates, den_rel):
    """The matched filter of several prepared curves ``(tau, yw, w, w_total, periods)``
    in one upload, one launch and one download."""
    ctx = _lib.context(0)
    device = _device_of(ctx)
    torch = _torch()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    sizes = [int(np.asarray(c[4]).size) for c in curves]
    n_out = sum(sizes)
    tau_d = dev(np.concatenate([c[0] for c in curves]))
    yw_d = dev(np.concatenate([c[1] for c in curves]))
    w_d = dev(np.concatenate([c[2] for c in curves]))
    per_d = dev(np.concatenate([np.asarray(c[4], dtype=np.float64) for c in curves]))
    desc = (_lib.TlsCurve * len(curves))()
    point = col = 0
    for i, (tau, _, _, w_total, _) in enumerate(curves):
        desc[i].point_offset = point
        desc[i].n_points = int(tau.size)
        desc[i].period_offset = col
        desc[i].n_periods = sizes[i]
        desc[i].out_offset = col
        desc[i].w_total = float(w_total)
        point += int(tau.size)
        col += sizes[i]
    widths = (ctypes.c_int32 * len(dur_bins))(*dur_bins)
    flat = np.ascontiguousarray(np.concatenate([templates[wd] for wd in dur_bins]),
                                dtype=np.float64)
    tmpl = flat.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    out = torch.empty((_lib.N_TLS_OUTPUTS, max(n_out, 1)), dtype=torch.float64, device=device)
    with ctx.lock:
        ctx.check(ctx.lib.cpb_tls_search(
            ctx.handle, _stream_handle(device), _ptr(tau_d), _ptr(yw_d), _ptr(w_d), point,
            desc, len(curves), _ptr(per_d), int(n_bins), widths, len(dur_bins), tmpl,
            float(den_rel), _ptr(out), n_out), "TLS search")
        host = out.cpu().numpy()
    results, col = [], 0
    for n in sizes:
        sl = slice(col, col + n)
        results.append({"sr": host[0, sl].copy(), "depth": host[1, sl].copy(),
                        "duration": host[2, sl].copy(), "t0": host[3, sl].copy()})
        col += n
    return results


def _prepare(t, y, dy):
    """``(t_ref, scale, tau, yw, w, w_total)`` of one curve, or ``None`` for a constant
    flux (src/cuperiod/methods/tls.py:676-693)."""
    t_ref = float(t.min())
    if _is_constant(y):
        return None
    scale = _flux_scale(y, dy)
    y = y / scale
    w = (np.ones_like(y) if dy is None
         else 1.0 / (np.ascontiguousarray(dy, dtype=np.float64) / scale) ** 2)
    w_total = float(w.sum())
    y_mean = float(np.dot(w, y) / w_total)
    return t_ref, scale, t - t_ref, w * (y - y_mean), w, w_total


def _finish(res, *, scale, t_ref, errors):
    """SDE, flux units and the absolute epoch (src/cuperiod/methods/tls.py:709-727)."""
    sr = res["sr"]
    mean, std = float(sr.mean()), float(sr.std())
    res["sde"] = (sr - mean) / std if std > 0.0 else np.zeros_like(sr)
    res["depth"] = res["depth"] * scale
    if not errors:
        res["sr"] = sr * (scale * scale)
    res["t0"] = res["t0"] + t_ref
    return res


def _templates(settings: TLSSettings):
    dur_bins = _duration_bins(settings)
    templates = {w: limb_darkened_template(w, settings.limb_dark_u1, settings.limb_dark_u2)
                 for w in dur_bins}
    if settings.n_phase_bins > 4096:
        raise BackendUnavailableError("TLS: more than 4096 phase bins per period")
    return dur_bins, templates


def tls_power_batch(
    curves: list[tuple[np.ndarray, np.ndarray, np.ndarray | None]],
    periods: list[np.ndarray],
    *,
    settings: TLSSettings,
    backend: str = HIP_BACKEND,
) -> list[dict[str, np.ndarray]]:
    """:func:`tls_power` of several ``(t, y, dy)`` curves, each over its own periods, in
    one engine call; every result equals that curve's :func:`tls_power`."""
    if backend != HIP_BACKEND:
        raise ValueError(f"unknown backend {backend!r}; this engine provides {HIP_BACKEND!r}")
    dur_bins, templates = _templates(settings)
    out: list[dict[str, np.ndarray] | None] = [None] * len(curves)
    live, meta = [], []
    for i, ((t, y, dy), per) in enumerate(zip(curves, periods, strict=True)):
        t = np.ascontiguousarray(t, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        per = np.ascontiguousarray(per, dtype=np.float64)
        if per.size == 0:
            out[i] = {key: np.zeros(0, dtype=np.float64) for key in _OUTPUTS}
            continue
        prep = _prepare(t, y, dy)
        if prep is None:
            out[i] = _no_transit(int(per.size), float(t.min()))
            continue
        t_ref, scale, tau, yw, w, w_total = prep
        live.append((tau, yw, w, w_total, per))
        meta.append((i, scale, t_ref, dy is not None))
    if live:
        found = _tls_hip_curves(live, n_bins=settings.n_phase_bins, dur_bins=dur_bins,
                                templates=templates, den_rel=_DEN_REL_EPS)
        for (i, scale, t_ref, errors), res in zip(meta, found):
            out[i] = _finish(res, scale=scale, t_ref=t_ref, errors=errors)
    return out  # type: ignore[return-value]


def tls_power(
    t: np.ndarray,
    y: np.ndarray,
    dy: np.ndarray | None,
    periods: np.ndarray,
    *,
    settings: TLSSettings,
    backend: str = HIP_BACKEND,
) -> dict[str, np.ndarray]:
    """TLS matched-filter search over ``periods`` (flux input; a transit is a dip),
    src/cuperiod/methods/tls.py:598-727.

    Returns float64 arrays aligned with ``periods``: ``sde``, ``sr`` (the chi-square
    reduction of the best level + transit fit; in flux units squared when ``dy`` is
    None), ``depth`` (flux units), ``duration`` (days) and ``t0`` in
    ``[t.min(), t.min() + period)``; where no dip qualifies ``depth`` and ``duration``
    are 0 and ``t0`` is ``t.min()``. A constant flux finds none at any period.
    """
    if backend != HIP_BACKEND:
        raise ValueError(f"unknown backend {backend!r}; this engine provides {HIP_BACKEND!r}")
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    periods_host = np.ascontiguousarray(periods, dtype=np.float64)
    n_periods = int(periods_host.size)
    if n_periods == 0:
        return {key: np.zeros(0, dtype=np.float64) for key in _OUTPUTS}
    prep = _prepare(t, y, dy)
    if prep is None:
        return _no_transit(n_periods, float(t.min()))
    t_ref, scale, tau, yw, w, w_total = prep
    dur_bins, templates = _templates(settings)
    if not dur_bins:
        return _no_transit(n_periods, t_ref)
    res = _tls_hip(tau, yw, w, periods_host, n_bins=settings.n_phase_bins,
                   dur_bins=dur_bins, templates=templates, w_total=w_total,
                   den_rel=_DEN_REL_EPS)
    return _finish(res, scale=scale, t_ref=t_ref, errors=dy is not None)


def tls_periodogram(finite: LightCurve, periods: np.ndarray, res: dict[str, np.ndarray],
                    backend: str) -> Periodogram:
    """The periodogram of one searched curve: SDE as the power, ``sr``, ``depth``,
    ``duration`` and ``t0`` as extras (src/cuperiod/methods/tls.py:799-823)."""
    return Periodogram.from_spectrum(
        method="TLS",
        backend=backend,
        frequency=1.0 / periods,
        power=res["sde"],
        objective_sense="max",
        n_samples=finite.n,
        baseline=finite.baseline,
        extras={"sr": res["sr"], "depth": res["depth"], "duration": res["duration"],
                "t0": res["t0"]},
        meta=finite.meta,
    )


class TLSMethod(PeriodogramMethod):
    """Transit Least Squares -- limb-darkened matched filter (backend ``"hip"``)."""

    name: ClassVar[str] = "TLS"
    objective_sense: ClassVar[str] = "max"
    supports_multiband: ClassVar[bool] = False
    natural_domain: ClassVar[Domain] = Domain.FLUX
    settings_cls: ClassVar[type] = TLSSettings
    cpu_backend: ClassVar[str | None] = None
    fast_cpu_backend: ClassVar[str | None] = None
    gpu_backend: ClassVar[str | None] = HIP_BACKEND
    portable_gpu_backend: ClassVar[str | None] = None
    all_backends: ClassVar[tuple[str, ...]] = (HIP_BACKEND,)

    def default_grid(self, lc: LightCurve, settings: TLSSettings) -> GridSpec:
        finite = lc.finite()
        periods = _period_grid(finite.baseline, settings, time=finite.time)
        if periods.size == 0:
            raise InsufficientDataError("TLS: baseline too short for the configured period range")
        return GridSpec(kind="period", values=periods, uniform=False)

    def power(
        self,
        grid: GridSpec,
        lc: LightCurve,
        settings: TLSSettings,
        backend: str,
        engine: object | None = None,
    ) -> Periodogram:
        finite = searchable_points(lc, settings.min_detections, "TLS")
        found = tls_power(finite.time, finite.value, finite.error, grid.period,
                          settings=settings, backend=backend)
        return tls_periodogram(finite, grid.period, found, backend)

    def estimate_device_bytes(self, n_points: int) -> int:
        return 64 * 1024**2 + n_points * 8 * 6


register(TLSMethod())

__all__ = [
    "TLSMethod",
    "limb_darkened_template",
    "tls_periodogram",
    "tls_power",
    "tls_power_batch",
]
