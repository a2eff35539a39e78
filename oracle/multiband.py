"""FP64 CPU oracle of the multi-band BLS statistics -- TEST INFRASTRUCTURE, never the
product.

* :func:`joint_bands` -- ``joint_bands`` (src/cuperiod/multiband/_bls_joint.py:94-128):
  per-band median centring, ``ivar = 1/dy^2``, one common time origin.
* :func:`joint_search` -- ``joint_search`` (``_bls_joint.py:151-204``) on its numba
  branch (``_joint_numba``, ``:510-528``) over the C++ restatement of
  ``_numba_joint_kernel._kernel`` (``:419-503``) in ``bls_joint_oracle.cpp``, finished by
  :func:`finish` (``_finish``, ``:359-406``).
* :func:`independent_stack` -- ``_independent`` (src/cuperiod/multiband/bls_mb.py:240-275)
  over the single-band oracle.

Pinning (besides retyping): one band reduces the joint statistic to the single-band
search (``tests/test_multiband_bls_shared.py:178-190``: power = clip(depth_snr, 0)^2,
same transit time); a brute-force sweep with exact box sums on tiny inputs; the
objective is invariant under a common time shift and under a per-band flux offset; a
phase-scrambled band no longer adds (``tests/test_multiband_bls_shared.py:193-228``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from oracle import _dptr, _lib, bls_power, duration_bins, ivar_floor, objective_flag

JOINT_FIELDS = ("power", "depth", "depth_snr", "duration", "transit_time", "band_index")


@dataclass(frozen=True)
class JointBands:
    """``JointBands`` (src/cuperiod/multiband/_bls_joint.py:65-91)."""

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
        return int(self.starts.size - 1)


def joint_bands(times, values, errors) -> JointBands:
    """``joint_bands`` (src/cuperiod/multiband/_bls_joint.py:94-128)."""
    t_min = float(min(float(np.min(t)) for t in times))
    taus, yws, ivars, sy, si, fl = [], [], [], [], [], []
    for t, y, dy in zip(times, values, errors, strict=True):
        y_c = np.asarray(y, dtype=np.float64)
        y_c = y_c - float(np.median(y_c))
        iv = 1.0 / (np.asarray(dy, dtype=np.float64) ** 2)
        yw = y_c * iv
        taus.append(np.asarray(t, dtype=np.float64) - t_min)
        yws.append(yw)
        ivars.append(iv)
        sy.append(float(yw.sum()))
        si.append(float(iv.sum()))
        fl.append(ivar_floor(int(yw.size), si[-1]))
    starts = np.concatenate([[0], np.cumsum([t.size for t in taus])]).astype(np.int64)
    return JointBands(
        tau=np.ascontiguousarray(np.concatenate(taus)),
        yw=np.ascontiguousarray(np.concatenate(yws)),
        ivar=np.ascontiguousarray(np.concatenate(ivars)),
        starts=starts,
        sum_y=np.asarray(sy, dtype=np.float64),
        sum_ivar=np.asarray(si, dtype=np.float64),
        floor=np.asarray(fl, dtype=np.float64),
        t_min=t_min,
    )


def segment_spec(periods, durations, oversample: int):
    """``_segment`` (src/cuperiod/multiband/_bls_joint.py:138-144):
    ``(periods, bin_duration, dur_bins, width)``."""
    periods = np.ascontiguousarray(periods, dtype=np.float64)
    durations = np.ascontiguousarray(durations, dtype=np.float64)
    bd = float(durations.min()) / oversample
    db = duration_bins(durations, bd)
    width = int(np.ceil(float(periods.max()) / bd)) + oversample + 1
    return periods, bd, db, width


def joint_kernel(bands: JointBands, periods, bin_duration: float, dur_bins, oversample: int,
                 width: int, obj_flag: int, threads: int = 0):
    """``_numba_joint_kernel._kernel``: ``(best_n, best_d, found, y_in, ivar_in)``."""
    lib = _lib()
    fn = lib.oracle_bls_joint
    if not getattr(fn, "_typed", False):
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_int64)
        fn.restype = ctypes.c_int
        fn.argtypes = [
            dp, dp, dp, lp, ctypes.c_int32, dp, dp, dp, dp, ctypes.c_int64, lp,
            ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
            ctypes.c_int32, lp, lp, ctypes.POINTER(ctypes.c_uint8), dp, dp,
        ]
        fn._typed = True
    per = np.ascontiguousarray(periods, dtype=np.float64)
    n_p = int(per.size)
    k = bands.n_bands
    db = np.ascontiguousarray(dur_bins, dtype=np.int64)
    starts = np.ascontiguousarray(bands.starts, dtype=np.int64)
    floors = np.ascontiguousarray(bands.floor, dtype=np.float64)
    o_n = np.zeros(max(n_p, 1), dtype=np.int64)
    o_d = np.zeros(max(n_p, 1), dtype=np.int64)
    o_f = np.zeros(max(n_p, 1), dtype=np.uint8)
    o_y = np.zeros(max(k * n_p, 1), dtype=np.float64)
    o_i = np.zeros(max(k * n_p, 1), dtype=np.float64)
    lp = ctypes.POINTER(ctypes.c_int64)
    rc = fn(
        _dptr(bands.tau), _dptr(bands.yw), _dptr(bands.ivar), starts.ctypes.data_as(lp), k,
        _dptr(np.ascontiguousarray(bands.sum_y)), _dptr(np.ascontiguousarray(bands.sum_ivar)),
        _dptr(floors), _dptr(per), n_p, db.ctypes.data_as(lp), int(db.size),
        float(bin_duration), int(oversample), int(width), int(obj_flag), int(threads),
        o_n.ctypes.data_as(lp), o_d.ctypes.data_as(lp),
        o_f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dptr(o_y), _dptr(o_i),
    )
    if rc != 0:
        raise RuntimeError(f"oracle_bls_joint failed with status {rc}")
    return (o_n[:n_p], o_d[:n_p], o_f[:n_p].astype(bool),
            o_y[: k * n_p].reshape(k, n_p), o_i[: k * n_p].reshape(k, n_p))


def finish(bands: JointBands, y_in, iv_in, best_n, best_d, found, periods,
           bin_duration: float, floors) -> dict[str, np.ndarray]:
    """``_finish`` (src/cuperiod/multiband/_bls_joint.py:359-406)."""
    k_all = bands.n_bands
    periods = np.asarray(periods, dtype=np.float64)
    n_p = int(periods.size)
    snr = np.zeros((k_all, n_p))
    depth = np.zeros((k_all, n_p))
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(k_all):
            iv_out = float(bands.sum_ivar[k]) - iv_in[k]
            ok = (iv_in[k] >= floors[k]) & (iv_out >= floors[k])
            s_in = np.where(ok, iv_in[k], 1.0)
            s_out = np.where(ok, iv_out, 1.0)
            yin = y_in[k] / s_in
            yout = (float(bands.sum_y[k]) - y_in[k]) / s_out
            dip = ok & (yout >= yin)
            d = yout - yin
            depth[k] = np.where(dip, d, 0.0)
            snr[k] = np.where(dip, d / np.sqrt(1.0 / s_in + 1.0 / s_out), 0.0)
    power = np.zeros(n_p)
    for k in range(k_all):
        power = power + snr[k] * snr[k]
    band = np.argmax(snr, axis=0) if k_all else np.zeros(n_p, dtype=np.int64)
    duration = np.asarray(best_d, dtype=np.float64) * bin_duration
    t0 = np.remainder(np.asarray(best_n, dtype=np.float64) * bin_duration + 0.5 * duration,
                      periods) + bands.t_min
    cols = np.arange(n_p)
    return {
        "power": np.where(found, power, 0.0),
        "depth": np.where(found, depth[band, cols], 0.0),
        "depth_snr": np.where(found, np.sqrt(power), 0.0),
        "duration": np.where(found, duration, 0.0),
        "transit_time": np.where(found, t0, 0.0),
        "band_index": np.where(found, band.astype(np.float64), 0.0),
    }


def joint_search(bands: JointBands, segments, oversample: int, *, objective: str = "snr",
                 threads: int = 0) -> list[dict[str, np.ndarray]]:
    """``joint_search`` on the numba branch: per ``(periods, durations)`` segment, the
    :data:`JOINT_FIELDS` aligned with its periods."""
    obj = objective_flag(objective)
    out = []
    for periods, durations in segments:
        periods = np.asarray(periods, dtype=np.float64)
        if periods.size == 0:
            out.append({name: np.zeros(0) for name in JOINT_FIELDS})
            continue
        per, bd, db, width = segment_spec(periods, durations, oversample)
        bn, bdur, found, y_in, iv_in = joint_kernel(bands, per, bd, db, oversample, width,
                                                    obj, threads)
        out.append(finish(bands, list(y_in), list(iv_in), bn, bdur, found, per, bd,
                          [float(f) for f in bands.floor]))
    return out


def independent_stack(times, values, errors, segments, oversample: int, *,
                      objective: str = "snr", threads: int = 0):
    """``_independent`` (src/cuperiod/multiband/bls_mb.py:240-275) over the single-band
    oracle: ``(period, power, extras)`` in grid order."""
    per_band = [
        [bls_power(t, y, dy, p, d, oversample, objective=objective, threads=threads)
         for p, d in segments]
        for t, y, dy in zip(times, values, errors, strict=True)
    ]
    parts = {n: [] for n in ("period", "power", "depth", "depth_snr", "duration", "t0",
                             "band_index")}
    for s, (periods, _) in enumerate(segments):
        seg = [b[s] for b in per_band]
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
    joined = {n: np.concatenate(v) for n, v in parts.items()}
    return joined.pop("period"), joined.pop("power"), joined


__all__ = [
    "JOINT_FIELDS",
    "JointBands",
    "finish",
    "independent_stack",
    "joint_bands",
    "joint_kernel",
    "joint_search",
    "segment_spec",
]
