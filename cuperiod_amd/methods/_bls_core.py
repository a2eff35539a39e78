"""BLS core on the MI355X: ``bls_power`` / ``bls_power_segments`` over the HIP engine.

Same signatures, host semantics and results objects as cuPeriod's BLS core
(src/cuperiod/methods/_bls_core.py):

* :class:`BLSPower` -- the per-period outputs, float64, in input order (``:113-131``);
* ``_host_inputs`` -- flux centred on its median in FP64, ``ivar = 1/dy**2``
  (``:1043-1058``);
* ``_search_spec`` / ``_duration_bins`` -- ``bd = min(durations)/oversample``, box
  widths ``floor(d/bd + 0.5)`` deduplicated in first-seen order, ``width =
  ceil(max P/bd) + oversample + 1`` (``:134-149, 697-706``);
* ``_ivar_floor`` -- the relative admissibility floor (``:75-94``);
* ``bls_power`` (``:882-1028``) and ``bls_power_segments`` (``:1075-1130``).

The box search itself runs in ``libcuperiod_bls.so`` (``_bls_search_hip_specs``, the
counterpart of ``_bls_search_cuda_specs``, ``:741-840``): one upload of the light
curve, one launch sequence for every segment, one (7, n_periods) download. Backend name:
``"hip"``. There is no CPU path in this package.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Any

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.core.errors import BackendUnavailableError

HIP_BACKEND = "hip"

_F64_EPS = float(np.finfo(np.float64).eps)
_F32_EPS = float(np.finfo(np.float32).eps)

#: Trial periods per batch of the reference's array backends; kept for signature
#: compatibility (the HIP engine tiles periods itself).
DEFAULT_BATCH = 2048

#: Output rows of every search, in the order the engine writes them.
_OUTPUT_NAMES = (
    "power",
    "depth",
    "depth_err",
    "depth_snr",
    "duration",
    "transit_time",
    "log_likelihood",
)


def _ivar_floor(n_points: int, sum_ivar: float, *, float32: bool = False) -> float:
    """Smallest admissible inverse-variance sum on either side of a box:
    ``4 * N * eps * sum(ivar)`` (and ``8 * eps32 * sum(ivar)`` for FP32 sums), never
    below the smallest normal float64."""
    floor = 4.0 * n_points * _F64_EPS * sum_ivar
    if float32:
        floor = max(floor, 8.0 * _F32_EPS * sum_ivar)
    if not floor > 0.0:
        return float(np.finfo(np.float64).tiny)
    return floor


@dataclass(frozen=True, slots=True)
class BLSPower:
    """Per-period BLS maxima aligned with the input ``periods`` (float64).

    ``power`` is ``depth_snr`` for the ``"snr"`` objective and ``log_likelihood`` for
    ``"likelihood"``; ``transit_time`` is the mid-transit time in ``[t_min, t_min +
    period)``; a period with no admissible box has ``power = -inf`` and zeros.
    """

    period: np.ndarray
    power: np.ndarray
    depth: np.ndarray
    depth_err: np.ndarray
    depth_snr: np.ndarray
    duration: np.ndarray
    transit_time: np.ndarray
    log_likelihood: np.ndarray


def _duration_bins(durations: np.ndarray, bin_duration: float) -> list[int]:
    """Box widths in bins, rounded half away from zero, deduplicated in first-seen
    order, only widths >= 1."""
    seen: set[int] = set()
    out: list[int] = []
    for d in np.asarray(durations, dtype=np.float64).tolist():
        kd = math.floor(d / bin_duration + 0.5)
        if kd >= 1 and kd not in seen:
            seen.add(kd)
            out.append(kd)
    return out


@dataclass(frozen=True, slots=True)
class _SearchSpec:
    """One duration-homogeneous box search: a period block and its binning."""

    periods: np.ndarray
    bin_duration: float
    dur_bins: list[int]
    width: int


def _search_spec(periods: np.ndarray, durations: np.ndarray, oversample: int) -> _SearchSpec:
    periods_host = np.ascontiguousarray(periods, dtype=np.float64)
    durations_host = np.ascontiguousarray(durations, dtype=np.float64)
    bin_duration = float(durations_host.min()) / oversample
    dur_bins = _duration_bins(durations_host, bin_duration)
    max_n_bins = int(np.ceil(float(periods_host.max()) / bin_duration)) + oversample
    return _SearchSpec(periods_host, bin_duration, dur_bins, max_n_bins + 1)


def _host_inputs(t, y, dy) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """``(t, y - median(y), 1/dy**2)`` as contiguous float64 host arrays."""
    t_host = np.ascontiguousarray(t, dtype=np.float64)
    y_host = np.ascontiguousarray(y, dtype=np.float64)
    if y_host.size:
        y_host = y_host - float(np.median(y_host))
    ivar_host = 1.0 / (np.ascontiguousarray(dy, dtype=np.float64) ** 2)
    return t_host, y_host, ivar_host


def _objective_flag(objective: str) -> int:
    """0 = depth signal-to-noise, 1 = log-likelihood."""
    if objective not in ("snr", "likelihood"):
        raise ValueError("objective must be 'snr' or 'likelihood'")
    return 0 if objective == "snr" else 1


def _empty_output(n_periods: int) -> dict[str, np.ndarray]:
    out = {name: np.zeros(n_periods, dtype=np.float64) for name in _OUTPUT_NAMES}
    out["power"][...] = -np.inf
    return out


def _empty_power() -> BLSPower:
    return BLSPower(*(np.empty(0, dtype=np.float64) for _ in range(8)))


def _to_power(periods: np.ndarray, out: dict[str, np.ndarray]) -> BLSPower:
    return BLSPower(
        period=periods,
        power=np.ascontiguousarray(out["power"], dtype=np.float64),
        depth=np.ascontiguousarray(out["depth"], dtype=np.float64),
        depth_err=np.ascontiguousarray(out["depth_err"], dtype=np.float64),
        depth_snr=np.ascontiguousarray(out["depth_snr"], dtype=np.float64),
        duration=np.ascontiguousarray(out["duration"], dtype=np.float64),
        transit_time=np.ascontiguousarray(out["transit_time"], dtype=np.float64),
        log_likelihood=np.ascontiguousarray(out["log_likelihood"], dtype=np.float64),
    )


def _check_backend(backend: str) -> None:
    if backend != HIP_BACKEND:
        raise ValueError(
            f"unknown backend {backend!r}; this engine provides {HIP_BACKEND!r}"
        )


# --- device plumbing ---------------------------------------------------------------


def _torch():
    import torch

    return torch


def _device_of(ctx: _lib.Context):
    torch = _torch()
    if not torch.cuda.is_available():
        raise BackendUnavailableError("BLS: torch sees no HIP device")
    return torch.device("cuda", ctx.device)


def _stream_handle(device) -> ctypes.c_void_p:
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(tensor.data_ptr())


def _hip_curve(t, y, ivar, device, device_cache: dict[str, Any] | None) -> dict[str, Any]:
    """The light curve's device arrays and host sums, cached per light curve
    (``_cuda_curve``, src/cuperiod/methods/_bls_core.py:709-738)."""
    data = None
    if device_cache is not None:
        data = device_cache.get("hip")
        if data is not None and data["n"] != t.shape[0]:
            data = None
    if data is None:
        torch = _torch()
        t_min = float(t.min())
        yw_h = y * ivar
        data = {
            "n": int(t.shape[0]),
            "t_min": t_min,
            "tau": torch.from_numpy(np.ascontiguousarray(t - t_min)).to(device),
            "yw": torch.from_numpy(np.ascontiguousarray(yw_h)).to(device),
            "ivar": torch.from_numpy(np.ascontiguousarray(ivar)).to(device),
            "sum_y": float(yw_h.sum()),
            "sum_ivar": float(ivar.sum()),
        }
        if device_cache is not None:
            device_cache["hip"] = data
    return data


def _bls_search_hip_specs(
    t: np.ndarray,
    y: np.ndarray,
    ivar: np.ndarray,
    specs: list[_SearchSpec],
    *,
    oversample: int,
    obj_flag: int,
    device_cache: dict[str, Any] | None = None,
    accumulation: int = 0,
) -> list[dict[str, np.ndarray]]:
    """Every segment's box search for one light curve: one upload, one launch
    sequence, one download. Each segment's result is that of a separate search.
    ``accumulation`` selects the bins (0 exact fixed point, 1 compensated FP32, 2 FP32;
    the FP32 modes use the FP32 admissibility floor)."""
    ctx = _lib.context(0)
    device = _device_of(ctx)
    torch = _torch()
    sizes = [int(spec.periods.shape[0]) for spec in specs]
    total = sum(sizes)
    if total == 0:
        return [_empty_output(0) for _ in specs]
    data = _hip_curve(t, y, ivar, device, device_cache)
    periods_d = torch.from_numpy(
        np.ascontiguousarray(np.concatenate([spec.periods for spec in specs]))
    ).to(device)
    dur_all = [np.asarray(spec.dur_bins, dtype=np.int32) for spec in specs]
    dur_h = np.concatenate(dur_all) if dur_all else np.zeros(0, dtype=np.int32)
    dur_d = torch.from_numpy(np.ascontiguousarray(dur_h if dur_h.size else np.zeros(1, np.int32))).to(device)
    pieces = [(k, a, b, w) for k, spec in enumerate(specs)
              for a, b, w in (_width_ranges(spec.periods, spec.bin_duration, oversample)
                              or [(0, 0, int(spec.width))])]
    segs = (_lib.Segment * max(1, len(pieces)))()
    d_offs = np.concatenate([[0], np.cumsum([d.shape[0] for d in dur_all])]).astype(int)
    p_base = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    for j, (k, a, b, w) in enumerate(pieces):
        sg = segs[j]
        sg.bin_duration = float(specs[k].bin_duration)
        sg.period_offset = int(p_base[k] + a)
        sg.n_periods = b - a
        sg.out_offset = int(p_base[k] + a)
        sg.width = int(w)
        sg.dur_offset = int(d_offs[k])
        sg.n_dur = int(dur_all[k].shape[0])
        sg.curve = 0
        sg.mode = 0
        sg.group = k
    if int(max(spec.width for spec in specs)) > (1 << 30):
        raise BackendUnavailableError("BLS: phase-bin count above 2^30 per period")
    out = torch.empty((_lib.N_OUTPUTS, total), dtype=torch.float64, device=device)
    floor = _ivar_floor(data["n"], data["sum_ivar"], float32=accumulation != 0)
    with ctx.lock:
        ctx.check(ctx.lib.cpb_set_accumulation(ctx.handle, int(accumulation)), "accumulation")
        status = ctx.lib.cpb_bls_search(
            ctx.handle, _stream_handle(device), _ptr(data["tau"]), _ptr(data["yw"]),
            _ptr(data["ivar"]), data["n"], data["t_min"], floor, data["sum_y"],
            data["sum_ivar"], _ptr(periods_d),
            _ptr(dur_d), segs, len(pieces), int(oversample), int(obj_flag), _ptr(out),
            total,
        )
        ctx.check(status, "search")
        host = out.cpu().numpy()
    results: list[dict[str, np.ndarray]] = []
    p_off = 0
    for n_periods in sizes:
        sl = slice(p_off, p_off + n_periods)
        results.append({name: host[i, sl].copy() for i, name in enumerate(_OUTPUT_NAMES)})
        p_off += n_periods
    return results


#: Phase-bin spans (bins per period, padded) where the engine's launch classes change:
#: 128-thread workgroups (640), registers and LDS with 8 waves per SIMD (1,152), registers
#: and LDS sized for 7, 6, 5 and 4 workgroups per CU (1,344, 1,600, 1,920, 256 threads x 8
#: bins), LDS only (160 KiB); wider spans go to global memory. The engine classifies a
#: segment by its widest span; cutting segments at the same limits keeps each period in
#: the class with the most workgroups per CU its span allows.
_SPAN_LIMITS = (640, 1152, 1344, 1600, 1920, 2048, 8960)


def _width_of(p_max: float, bin_duration: float, oversample: int) -> int:
    """``_search_spec``'s histogram width for periods up to ``p_max``."""
    return int(math.ceil(p_max / bin_duration)) + int(oversample) + 1


def _width_ranges(periods: np.ndarray, bin_duration: float, oversample: int
                  ) -> list[tuple[int, int, int]]:
    """``[(start, stop, width)]``: a monotone period array cut where the width its
    periods need crosses the engine's span limits, so narrow periods keep the fast
    layouts. A period's outputs do not depend on the width (``n_bins = min(ceil(P/bd)
    + os, width - 1)`` never binds, and no bin index reaches ``width - 1``); unsorted
    arrays stay whole."""
    n = int(periods.size)
    whole = [(0, n, _width_of(float(periods.max()), bin_duration, oversample))] if n else []
    if n < 2:
        return whole
    d = np.diff(periods)
    if not (np.all(d > 0) or np.all(d < 0)):
        return whole
    if whole[0][2] <= _SPAN_LIMITS[0]:
        return whole
    need = np.ceil(periods / bin_duration).astype(np.int64) + int(oversample) + 1
    cls = np.searchsorted(np.asarray(_SPAN_LIMITS), need, side="left")
    cuts = np.flatnonzero(np.diff(cls)) + 1
    bounds = [0, *cuts.tolist(), n]
    return [(a, b, int(need[a:b].max())) for a, b in zip(bounds[:-1], bounds[1:])]


# --- public entry points -------------------------------------------------------------

#: Bin accumulation modes (``BLSSettings.accumulation``) and their engine codes.
ACCUMULATION_MODES = {"fixed64": 0, "kahan32": 1, "fp32": 2}


def _accumulation_code(accumulation: str | None, precision: str) -> int:
    """Engine code of an accumulation mode; ``precision="float32"`` (the reference's
    float32 kernel) means plain FP32 bins when no mode is named."""
    if accumulation is None:
        return 2 if str(precision).lower() in ("float32", "fp32", "single") else 0
    try:
        return ACCUMULATION_MODES[str(accumulation)]
    except KeyError:
        raise ValueError(
            f"unknown accumulation {accumulation!r}; choose one of "
            f"{sorted(ACCUMULATION_MODES)}"
        ) from None


def bls_power(
    t: np.ndarray,
    y: np.ndarray,
    dy: np.ndarray,
    periods: np.ndarray,
    durations: np.ndarray,
    oversample: int,
    *,
    objective: str = "snr",
    backend: str = HIP_BACKEND,
    batch: int = DEFAULT_BATCH,
    precision: str = "auto",
    device_cache: dict[str, Any] | None = None,
    accumulation: str | None = None,
) -> BLSPower:
    """BLS box search of one light curve over ``periods`` on the MI355X.

    Equivalent to ``BoxLeastSquares(t, y, dy).power(periods, durations,
    objective=objective, oversample=oversample)`` with cuPeriod's relative
    admissibility floor: median-centred flux, ``ivar = 1/dy**2``, box widths rounded
    half away from zero. ``batch`` is accepted for signature compatibility. The engine
    bins in exact 64-bit fixed point unless ``accumulation`` is ``"kahan32"`` or
    ``"fp32"`` (or ``precision="float32"``), the reduced-precision modes of the
    precision study.

    Raises
    ------
    ValueError
        For an unknown objective or backend.
    BackendUnavailableError
        Without the HIP library or a device.
    """
    obj_flag = _objective_flag(objective)
    _check_backend(backend)
    periods_host = np.ascontiguousarray(periods, dtype=np.float64)
    if periods_host.size == 0:
        return _empty_power()
    spec = _search_spec(periods_host, durations, oversample)
    t_host, y_host, ivar_host = _host_inputs(t, y, dy)
    if not spec.dur_bins:
        return _to_power(periods_host, _empty_output(periods_host.size))
    out = _bls_search_hip_specs(
        t_host, y_host, ivar_host, [spec], oversample=oversample, obj_flag=obj_flag,
        device_cache=device_cache, accumulation=_accumulation_code(accumulation, precision),
    )[0]
    return _to_power(periods_host, out)


def bls_power_segments(
    t: np.ndarray,
    y: np.ndarray,
    dy: np.ndarray,
    segments: list[tuple[np.ndarray, np.ndarray]],
    oversample: int,
    *,
    objective: str = "snr",
    backend: str = HIP_BACKEND,
    batch: int = DEFAULT_BATCH,
    precision: str = "auto",
    device_cache: dict[str, Any] | None = None,
    accumulation: str | None = None,
) -> list[BLSPower]:
    """:func:`bls_power` over the ``(periods, durations)`` segments of one curve,
    with one upload of the curve and one download of every result; equal to one
    :func:`bls_power` call per segment."""
    obj_flag = _objective_flag(objective)
    _check_backend(backend)
    acc = _accumulation_code(accumulation, precision)
    t_host, y_host, ivar_host = _host_inputs(t, y, dy)
    specs = [
        _search_spec(periods, durations, oversample)
        for periods, durations in segments
        if np.asarray(periods).size
    ]
    outs = iter(
        _bls_search_hip_specs(
            t_host, y_host, ivar_host, specs, oversample=oversample, obj_flag=obj_flag,
            device_cache=device_cache, accumulation=acc,
        )
        if specs
        else []
    )
    spec_iter = iter(specs)
    results: list[BLSPower] = []
    for periods, _ in segments:
        if np.asarray(periods).size == 0:
            results.append(_empty_power())
        else:
            results.append(_to_power(next(spec_iter).periods, next(outs)))
    return results


__all__ = [
    "ACCUMULATION_MODES",
    "BLSPower",
    "DEFAULT_BATCH",
    "HIP_BACKEND",
    "bls_power",
    "bls_power_segments",
]
