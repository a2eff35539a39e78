"""FP64 CPU oracle for the BLS hot path -- TEST INFRASTRUCTURE, never the product.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import
this package, and only as the checker. The product package (``cuperiod_amd``) never
imports it and has no CPU fallback through it.

Contents
--------
* :func:`bls_search` -- ``_bls_search_numba`` (src/cuperiod/methods/_bls_core.py:468-500)
  over the C++ restatement of ``_numba_bls_kernel`` in ``bls_oracle.cpp``
  (``_bls_core.py:334-465``), screened or unscreened.
* :func:`bls_power` -- the numba branch of ``bls_power`` (``_bls_core.py:882-1028``):
  ``_host_inputs`` (``:1043-1058``), ``_search_spec`` (``:697-706``),
  ``_duration_bins`` (``:134-149``), ``_ivar_floor`` (``:75-94``).
* :mod:`oracle.bruteforce` -- direct, exactly-summed box statistics for tiny inputs.

Pinning (what checks each function, besides retyping it):
* ``duration_bins``: the reference's known answer ``[0.2,0.25,0.45]/0.1 -> [2,3,5]``
  (tests/test_bls.py:213-215) and half-away rounding cases.
* ``bls_search``: the brute-force checker (exact sums, no prefix sums), a hand-built
  wrap-copy case, screened == unscreened bit for bit on the reference's 16 adversarial
  curves (tests/test_bls.py:374-424), offset/scale invariance (tests/test_bls.py:
  182-210), period/depth recovery on the reference's eclipser recipe (tests/synth.py:
  167-186, tests/test_bls.py:73-79) and the closed-form objectives (Kovacs et al. 2002).
* ``ivar_floor``: closed form plus the invariance checks above.
"""

from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle_bls.so"
_LIB: ctypes.CDLL | None = None

_F64_EPS = float(np.finfo(np.float64).eps)
_F32_EPS = float(np.finfo(np.float32).eps)
_SCREEN_SLACK_EPS = 256.0
_SCREEN_SCALE_RANGE = (1e-150, 1e150)

OUTPUT_NAMES = (
    "power",
    "depth",
    "depth_err",
    "depth_snr",
    "duration",
    "transit_time",
    "log_likelihood",
)


def build() -> Path:
    """Compile ``liboracle_bls.so`` with the committed Makefile (g++, OpenMP)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def _lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not _LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(_LIB_PATH))
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_int64)
        lib.oracle_bls_search.restype = ctypes.c_int
        lib.oracle_bls_search.argtypes = [
            dp, dp, dp, ctypes.c_int64, dp, ctypes.c_int64, lp, ctypes.c_int64,
            ctypes.c_double, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp,
        ]
        lib.oracle_max_threads.restype = ctypes.c_int
        lib.oracle_max_threads.argtypes = []
        _LIB = lib
    return _LIB


def max_threads() -> int:
    """OpenMP threads the oracle uses by default on this host."""
    return int(_lib().oracle_max_threads())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ivar_floor(n_points: int, sum_ivar: float, *, float32: bool = False) -> float:
    """``_ivar_floor`` (src/cuperiod/methods/_bls_core.py:75-94)."""
    floor = 4.0 * n_points * _F64_EPS * sum_ivar
    if float32:
        floor = max(floor, 8.0 * _F32_EPS * sum_ivar)
    if not floor > 0.0:
        return float(np.finfo(np.float64).tiny)
    return floor


def duration_bins(durations, bin_duration: float) -> list[int]:
    """``_duration_bins`` (src/cuperiod/methods/_bls_core.py:134-149)."""
    seen: set[int] = set()
    out: list[int] = []
    for d in np.asarray(durations, dtype=np.float64).tolist():
        kd = math.floor(d / bin_duration + 0.5)
        if kd >= 1 and kd not in seen:
            seen.add(kd)
            out.append(kd)
    return out


def search_spec(periods, durations, oversample: int):
    """``_search_spec`` (src/cuperiod/methods/_bls_core.py:697-706).

    Returns ``(periods, bin_duration, dur_bins, width)``.
    """
    periods_host = np.ascontiguousarray(periods, dtype=np.float64)
    durations_host = np.ascontiguousarray(durations, dtype=np.float64)
    bin_duration = float(durations_host.min()) / oversample
    dur_bins = duration_bins(durations_host, bin_duration)
    max_n_bins = int(np.ceil(float(periods_host.max()) / bin_duration)) + oversample
    return periods_host, bin_duration, dur_bins, max_n_bins + 1


def host_inputs(t, y, dy):
    """``_host_inputs`` (src/cuperiod/methods/_bls_core.py:1043-1058)."""
    t_host = np.ascontiguousarray(t, dtype=np.float64)
    y_host = np.ascontiguousarray(y, dtype=np.float64)
    if y_host.size:
        y_host = y_host - float(np.median(y_host))
    ivar_host = 1.0 / (np.ascontiguousarray(dy, dtype=np.float64) ** 2)
    return t_host, y_host, ivar_host


def bls_search(
    t, y, ivar, periods, *, bin_duration: float, dur_bins, oversample: int,
    width: int, obj_flag: int, screen: bool = True, threads: int = 0,
) -> dict[str, np.ndarray]:
    """``_bls_search_numba`` (src/cuperiod/methods/_bls_core.py:468-500).

    ``t`` absolute times, ``y`` centred flux, ``ivar`` inverse variances.
    """
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    ivar = np.asarray(ivar, dtype=np.float64)
    yw = np.ascontiguousarray(y * ivar, dtype=np.float64)
    t_min = float(t.min())
    tau = np.ascontiguousarray(t - t_min, dtype=np.float64)
    db = np.ascontiguousarray(dur_bins, dtype=np.int64)
    sum_y = float(yw.sum())
    sum_ivar = float(ivar.sum())
    sum_abs = float(np.abs(yw).sum())
    lo, hi = _SCREEN_SCALE_RANGE
    if screen and lo <= sum_abs <= hi and lo <= sum_ivar <= hi:
        inv_a, inv_s = 1.0 / sum_abs, 1.0 / sum_ivar
        slack = _SCREEN_SLACK_EPS * _F64_EPS
    else:
        inv_a, inv_s, slack = 0.0, 0.0, math.inf
    wv = np.ascontiguousarray(ivar, dtype=np.float64)
    per = np.ascontiguousarray(periods, dtype=np.float64)
    n_periods = int(per.size)
    out = np.empty(7 * max(n_periods, 1), dtype=np.float64)
    rc = _lib().oracle_bls_search(
        _dptr(tau), _dptr(yw), _dptr(wv), int(tau.size), _dptr(per), n_periods,
        db.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(db.size),
        float(bin_duration), int(oversample), int(width), sum_y, sum_ivar, t_min,
        int(obj_flag), ivar_floor(int(tau.size), sum_ivar), inv_a, inv_s, slack,
        int(threads), _dptr(out),
    )
    if rc != 0:
        raise RuntimeError(f"oracle_bls_search failed with status {rc}")
    out = out[: 7 * n_periods].reshape(7, n_periods)
    return {name: out[i].copy() for i, name in enumerate(OUTPUT_NAMES)}


def objective_flag(objective: str) -> int:
    if objective not in ("snr", "likelihood"):
        raise ValueError("objective must be 'snr' or 'likelihood'")
    return 0 if objective == "snr" else 1


def bls_power(
    t, y, dy, periods, durations, oversample: int, *, objective: str = "snr",
    screen: bool = True, threads: int = 0,
) -> dict[str, np.ndarray]:
    """The numba branch of ``bls_power`` (src/cuperiod/methods/_bls_core.py:882-1028).

    Returns a dict with ``period`` plus the seven outputs, aligned with ``periods``.
    """
    obj_flag = objective_flag(objective)
    periods_host = np.ascontiguousarray(periods, dtype=np.float64)
    if periods_host.size == 0:
        return {"period": periods_host, **{k: np.empty(0) for k in OUTPUT_NAMES}}
    per, bd, dur_bins, width = search_spec(periods_host, durations, oversample)
    t_host, y_host, ivar_host = host_inputs(t, y, dy)
    out = bls_search(
        t_host, y_host, ivar_host, per, bin_duration=bd, dur_bins=dur_bins,
        oversample=oversample, width=width, obj_flag=obj_flag, screen=screen,
        threads=threads,
    )
    return {"period": per, **out}


def default_threads() -> int:
    """Threads for a timed CPU baseline: every core the host exposes."""
    return int(os.environ.get("ORACLE_THREADS", "0")) or max_threads()


__all__ = [
    "OUTPUT_NAMES",
    "bls_power",
    "bls_search",
    "build",
    "duration_bins",
    "host_inputs",
    "ivar_floor",
    "max_threads",
    "objective_flag",
    "search_spec",
]
