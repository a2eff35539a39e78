"""Joint-ephemeris multi-band box search on the MI355X engine.

A transit is achromatic: every band dips at the same mid-transit time for the same
duration, only the depth changes with wavelength. The joint statistic places one box
-- one ``(duration, start)`` on a phase grid common to all bands, binned from one time
origin -- and lets each band fit its own depth inside it::

    J(box) = sum_k max(depth_k, 0)^2 / (1/ivar_in_k + 1/ivar_out_k)

(``objective="likelihood"``: the summed ``0.5*ivar_in_k*depth_k^2``), maximised over the
boxes of each trial period. The reported power is ``sum_k max(SNR_k, 0)^2`` at the chosen
box, its geometry the shared box, its depth that of the band with the largest SNR.

Host preparation restates ``joint_bands`` (src/cuperiod/multiband/_bls_joint.py:94-128)
and ``_segment`` (``:138-144``); the sweep and the finish run in one HIP launch
(``cpb_bls_joint_search``), replacing ``_joint_cuda`` (``:703-796``) and ``_finish``
(``:359-406``). Per-band sums are exact 64-bit fixed point, so the result does not depend
on the order of the points; ``precision`` and ``batch`` are accepted for the reference's
signature and do not change the search.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Final

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.core.errors import BackendUnavailableError
from cuperiod_amd.methods._bls_core import (
    _check_backend,
    _device_of,
    _ivar_floor,
    _objective_flag,
    _ptr,
    _search_spec,
    _stream_handle,
    _torch,
)

#: Fields of one joint search, aligned with the trial periods.
JOINT_FIELDS: Final = ("power", "depth", "depth_snr", "duration", "transit_time", "band_index")

#: Bands one star may hold in a joint search.
MAX_BANDS: Final = 64


@dataclass(frozen=True)
class JointBands:
    """Every contributing band of one star, ready for the joint sweep (host float64).

    ``tau`` holds each band's times minus the common origin ``t_min``; ``yw``/``ivar``
    the median-centred ``y*ivar`` and ``1/dy**2``; band ``k`` occupies
    ``starts[k]:starts[k + 1]`` of the concatenated arrays; ``sum_y``/``sum_ivar``/
    ``floor`` are per band.
    """

    tau: np.ndarray
    yw: np.ndarray
    ivar: np.ndarray
    starts: np.ndarray
    sum_y: np.ndarray
    sum_ivar: np.ndarray
    floor: np.ndarray
    t_min: float

    @property
    def n_bands(self) -> int:
        """Number of bands."""
        return len(self.starts) - 1

    def band(self, k: int) -> slice:
        """Where band ``k`` lies in the concatenated arrays."""
        lo, hi = self.starts[k:k + 2]
        return slice(int(lo), int(hi))


def _centred_band(t, y, dy, t_min: float) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """One band's ``(t - t_min, (y - median(y)) / dy^2, 1 / dy^2)``."""
    y = np.asarray(y, dtype=np.float64)
    weight = 1.0 / np.asarray(dy, dtype=np.float64) ** 2
    return np.asarray(t, dtype=np.float64) - t_min, (y - float(np.median(y))) * weight, weight


def joint_bands(times: list[np.ndarray], values: list[np.ndarray],
                errors: list[np.ndarray]) -> JointBands:
    """Prepare the bands of one star: each centred on its own median, ``ivar =
    1/dy^2``, times from the earliest time of any band."""
    t_min = min(float(np.min(t)) for t in times)
    prepared = [_centred_band(t, y, dy, t_min)
                for t, y, dy in zip(times, values, errors, strict=True)]
    tau, yw, ivar = (np.concatenate(column) for column in zip(*prepared))
    sizes = np.array([band[0].size for band in prepared], dtype=np.int64)
    sum_y = np.array([float(band[1].sum()) for band in prepared], dtype=np.float64)
    sum_ivar = np.array([float(band[2].sum()) for band in prepared], dtype=np.float64)
    floor = np.array([_ivar_floor(int(n), float(w)) for n, w in zip(sizes, sum_ivar)],
                     dtype=np.float64)
    return JointBands(tau=tau, yw=yw, ivar=ivar,
                      starts=np.concatenate([np.zeros(1, dtype=np.int64), np.cumsum(sizes)]),
                      sum_y=sum_y, sum_ivar=sum_ivar, floor=floor, t_min=t_min)


def _empty(n: int) -> dict[str, np.ndarray]:
    return {name: np.zeros(n, dtype=np.float64) for name in JOINT_FIELDS}


def joint_search(
    bands: JointBands,
    segments: list[tuple[np.ndarray, np.ndarray]],
    oversample: int,
    *,
    objective: str,
    backend: str,
    batch: int = 2048,
    precision: str = "auto",
) -> list[dict[str, np.ndarray]]:
    """The joint-ephemeris box search over every ``(periods, durations)`` segment.

    Returns, per segment, :data:`JOINT_FIELDS` arrays aligned with its periods (host
    float64); a period without any admissible box has power 0 and zero geometry.
    """
    return joint_search_stars([bands], [segments], oversample, objective=objective,
                              backend=backend)[0]


def joint_search_stars(
    stars: list[JointBands],
    segments: list[list[tuple[np.ndarray, np.ndarray]]],
    oversample: int,
    *,
    objective: str,
    backend: str,
) -> list[list[dict[str, np.ndarray]]]:
    """:func:`joint_search` of several stars (each with its own segments) in one
    upload, one launch and one download; results per star, per segment."""
    _check_backend(backend)
    obj_flag = _objective_flag(objective)
    for b in stars:
        if b.n_bands > MAX_BANDS:
            raise BackendUnavailableError(
                f"BLS multiband: {b.n_bands} bands; the joint search holds at most "
                f"{MAX_BANDS} per star"
            )
    specs = [
        [_search_spec(np.asarray(p, dtype=np.float64), np.asarray(d, dtype=np.float64),
                      oversample) if np.asarray(p).size else None for p, d in segs]
        for segs in segments
    ]
    live = [[s for s in row if s is not None] for row in specs]
    found = _joint_hip(stars, live, oversample, obj_flag) if any(live) else [[] for _ in stars]
    out = []
    for row, res in zip(specs, found):
        it = iter(res)
        out.append([_empty(0) if s is None else next(it) for s in row])
    return out


def _joint_hip(stars: list[JointBands], specs: list[list], oversample: int, obj_flag: int
               ) -> list[list[dict[str, np.ndarray]]]:
    """One upload, one launch, one download for every segment of every star."""
    ctx = _lib.context(0)
    device = _device_of(ctx)
    torch = _torch()
    flat = [(s, spec) for s, row in enumerate(specs) for spec in row]
    if max(int(spec.width) for _, spec in flat) > (1 << 24):
        raise BackendUnavailableError("BLS multiband: phase-bin count above 2^24 per period")
    sizes = [int(spec.periods.shape[0]) for _, spec in flat]
    total = sum(sizes)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    tau = dev(np.concatenate([b.tau for b in stars]))
    yw = dev(np.concatenate([b.yw for b in stars]))
    ivar = dev(np.concatenate([b.ivar for b in stars]))
    periods = dev(np.concatenate([spec.periods for _, spec in flat]))
    dur_all = [np.asarray(spec.dur_bins, dtype=np.int32) for _, spec in flat]
    dur_h = np.concatenate(dur_all)
    dur = dev(dur_h if dur_h.size else np.zeros(1, np.int32))
    n_bands = sum(b.n_bands for b in stars)
    band_desc = (_lib.Band * n_bands)()
    star_desc = (_lib.Star * len(stars))()
    point = band = 0
    for s, b in enumerate(stars):
        star_desc[s].band_offset = band
        star_desc[s].n_bands = b.n_bands
        star_desc[s].t_min = float(b.t_min)
        for k in range(b.n_bands):
            band_desc[band].point_offset = point + int(b.starts[k])
            band_desc[band].n_points = int(b.starts[k + 1] - b.starts[k])
            band_desc[band].ivar_floor = float(b.floor[k])
            band += 1
        point += int(b.tau.size)
    segs = (_lib.Segment * len(flat))()
    d_off = np.concatenate([[0], np.cumsum([d.shape[0] for d in dur_all])]).astype(int)
    p_off = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    for j, (s, spec) in enumerate(flat):
        sg = segs[j]
        sg.bin_duration = float(spec.bin_duration)
        sg.period_offset = int(p_off[j])
        sg.n_periods = sizes[j]
        sg.out_offset = int(p_off[j])
        sg.width = int(spec.width)
        sg.dur_offset = int(d_off[j])
        sg.n_dur = int(dur_all[j].shape[0])
        sg.curve = s
        sg.mode = 0
    out = torch.empty((_lib.N_JOINT_OUTPUTS, total), dtype=torch.float64, device=device)
    with ctx.lock:
        status = ctx.lib.cpb_bls_joint_search(
            ctx.handle, _stream_handle(device), _ptr(tau), _ptr(yw), _ptr(ivar), point,
            band_desc, n_bands, star_desc, len(stars), _ptr(periods), _ptr(dur), segs,
            len(flat), int(oversample), int(obj_flag), _ptr(out), total,
        )
        ctx.check(status, "joint search")
        host = out.cpu().numpy()
    results: list[list[dict[str, np.ndarray]]] = [[] for _ in stars]
    for j, (s, _) in enumerate(flat):
        sl = slice(int(p_off[j]), int(p_off[j + 1]))
        results[s].append({name: host[i, sl].copy() for i, name in enumerate(JOINT_FIELDS)})
    return results


__all__ = [
    "JOINT_FIELDS",
    "MAX_BANDS",
    "JointBands",
    "joint_bands",
    "joint_search",
    "joint_search_stars",
]
