"""Peak candidates and N-best selection on a score to maximise.

The selection rules of cuPeriod's peak picker (src/cuperiod/core/peaks.py:57-231):
interior local maxima (strictly above the left neighbour, at least the right one, so a
flat top reports its left edge), greedy selection by descending score with a minimum
frequency separation, and the alias-diverse variant that also skips harmonics,
sub-harmonics and the sidereal/solar/yearly alias lattice of every chosen peak.
"""

from __future__ import annotations

from typing import Literal

import numpy as np

#: Sidereal-day frequency (cycles/day): the dominant ground-based alias spacing.
SIDEREAL_FREQ_CPD = 1.0027379
#: Solar-day frequency (cycles/day).
SOLAR_FREQ_CPD = 1.0
#: Yearly sidelobe spacing (cycles/day).
YEARLY_FREQ_CPD = 1.0 / 365.25

def _sum_is_finite(values: np.ndarray) -> bool:
    with np.errstate(invalid="ignore", over="ignore"):
        return bool(np.isfinite(values.sum()))


def _finite_or_minus_inf(score: np.ndarray) -> np.ndarray:
    if _sum_is_finite(score):
        return score
    return np.where(np.isfinite(score), score, -np.inf)


def _interior_maxima(safe: np.ndarray) -> np.ndarray:
    if safe.size < 3:
        return np.empty(0, dtype=np.int64)
    interior = (safe[1:-1] > safe[:-2]) & (safe[1:-1] >= safe[2:])
    return np.flatnonzero(interior).astype(np.int64) + 1


def local_maxima(score: np.ndarray) -> np.ndarray:
    """Indices of interior local maxima (non-finite samples compare as ``-inf``)."""
    if score.size < 3:
        return np.empty(0, dtype=np.int64)
    return _interior_maxima(_finite_or_minus_inf(score))


def _descending(score: np.ndarray, candidates: np.ndarray) -> np.ndarray:
    """``candidates`` by descending score; equal scores in reverse index order, as
    ``np.argsort(score)[::-1]`` orders them."""
    return candidates[np.argsort(score[candidates])[::-1]]


def select_top_peaks(
    frequency: np.ndarray,
    score: np.ndarray,
    candidates: np.ndarray,
    n_peaks: int,
    min_separation: float,
    *,
    examined: list | None = None,
) -> np.ndarray:
    """Greedy highest-score peaks at least ``min_separation`` apart in frequency.
    ``examined`` (a list) receives the number of candidates visited."""
    if n_peaks <= 0:
        return np.empty(0, dtype=np.int64)
    chosen: list[int] = []
    seen = 0
    for idx in _descending(score, candidates):
        seen += 1
        if all(abs(frequency[idx] - frequency[j]) >= min_separation for j in chosen):
            chosen.append(int(idx))
            if len(chosen) >= n_peaks:
                break
    if examined is not None:
        examined.append(seen)
    return np.asarray(chosen, dtype=np.int64)


def _alias_lattice(harmonic_max: int) -> tuple[np.ndarray, np.ndarray]:
    """Harmonic orders ``2..harmonic_max`` and the alias offsets (sidereal and solar day
    at k = 1, 2, then the yearly sidelobe), as float64 vectors."""
    orders = np.arange(2, harmonic_max + 1, dtype=np.float64)
    day = np.array([SIDEREAL_FREQ_CPD, SOLAR_FREQ_CPD], dtype=np.float64)
    offsets = np.concatenate([(day[:, None] * np.array([1.0, 2.0])[None, :]).ravel(),
                              [YEARLY_FREQ_CPD]])
    return orders, offsets


def _related(f_new, f_chosen, tol: float, separation: float, orders: np.ndarray,
             offsets: np.ndarray) -> np.ndarray:
    """The test of :func:`is_alias_related` element-wise over ``f_new`` and ``f_chosen``
    (arrays or scalars, broadcast against each other; the harmonic orders and alias
    offsets run along an extra last axis): does ``f_new`` duplicate, harmonise or alias
    ``f_chosen``."""
    a = np.asarray(f_new, dtype=np.float64)[..., None]
    b = np.asarray(f_chosen, dtype=np.float64)[..., None]
    near = (np.abs(a - b) < separation)[..., 0]
    harmonic = (np.abs(a - orders * b) < orders * tol).any(-1)
    subharmonic = (np.abs(a - b / orders) < tol).any(-1)
    below = (np.abs(a - np.abs(b - offsets)) < tol).any(-1)
    above = (np.abs(a - (b + offsets)) < tol).any(-1)
    return near | harmonic | subharmonic | below | above


def is_alias_related(
    f_new: float,
    f_chosen: float,
    tol: float,
    harmonic_max: int,
    *,
    min_separation: float | None = None,
) -> bool:
    """Whether ``f_new`` duplicates ``f_chosen`` (closer than ``min_separation``, default
    ``tol``), is one of its harmonics / sub-harmonics up to ``harmonic_max`` (within
    ``m*tol`` of the m-th harmonic, ``tol`` of a sub-harmonic), or sits within ``tol`` of
    a sidereal/solar (k = 1, 2) or yearly alias on either side of it."""
    orders, offsets = _alias_lattice(harmonic_max)
    sep = tol if min_separation is None else min_separation
    return bool(_related(float(f_new), float(f_chosen), tol, sep, orders, offsets))


_NATIVE: list = []


def _native_library():
    """The engine library when it loads (its host-side selector is used then), else
    ``None`` and the selection runs in numpy."""
    if not _NATIVE:
        try:
            from cuperiod_amd import _lib

            _NATIVE.append(_lib.load_library())
        except Exception:  # noqa: BLE001 - any load failure: the numpy rule
            _NATIVE.append(None)
    return _NATIVE[0]


def _select_native(lib, frequency, order, n_peaks, tol, sep, orders, offsets, examined):
    import ctypes

    freq = np.ascontiguousarray(frequency, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int64)
    orders = np.ascontiguousarray(orders, dtype=np.float64)
    offsets = np.ascontiguousarray(offsets, dtype=np.float64)
    chosen = np.empty(max(int(n_peaks), 1), dtype=np.int64)
    count, visited = ctypes.c_int32(), ctypes.c_int64()
    status = lib.cpb_select_diverse_peaks(
        freq.ctypes.data, order.ctypes.data, order.size, int(n_peaks), float(tol),
        float(sep), orders.ctypes.data, orders.size, offsets.ctypes.data, offsets.size,
        chosen.ctypes.data, ctypes.byref(count), ctypes.byref(visited))
    if status != 0:
        raise ValueError(f"alias-diverse selection failed (status {status})")
    if examined is not None:
        examined.append(int(visited.value))
    return chosen[:count.value].copy()


def select_diverse_peaks(
    frequency: np.ndarray,
    score: np.ndarray,
    candidates: np.ndarray,
    n_peaks: int,
    tol: float,
    harmonic_max: int,
    *,
    min_separation: float | None = None,
    examined: list | None = None,
) -> np.ndarray:
    """Greedy highest-score peaks skipping any candidate alias-related (either way
    round) to an already chosen one. ``examined`` (a list) receives the number of
    candidates visited.

    Candidates are taken in score order a block at a time: the block is tested against
    every peak chosen so far in one vectorised pass, and each peak chosen inside it
    blocks the rest of the block the same way; the first unblocked candidate is always
    the greedy rule's next choice."""
    if n_peaks <= 0:
        return np.empty(0, dtype=np.int64)
    orders, offsets = _alias_lattice(harmonic_max)
    sep = tol if min_separation is None else min_separation
    order = _descending(score, candidates)
    lib = _native_library()
    if lib is not None:
        return _select_native(lib, frequency, order, n_peaks, tol, sep, orders, offsets,
                              examined)

    def related(a, b):
        return (_related(a, b, tol, sep, orders, offsets)
                | _related(b, a, tol, sep, orders, offsets))

    f = np.asarray(frequency[order], dtype=np.float64)
    picked: list[int] = []
    visited = order.size
    block = 128
    for lo in range(0, order.size, block):
        fb = f[lo:lo + block]
        if picked:
            blocked = related(fb[None, :], f[picked][:, None]).any(0)
        else:
            blocked = np.zeros(fb.size, dtype=bool)
        j = 0
        while True:
            free = np.flatnonzero(~blocked[j:])
            if free.size == 0:
                break
            j += int(free[0])
            picked.append(lo + j)
            if len(picked) >= n_peaks:
                visited = lo + j + 1
                break
            blocked[j + 1:] |= related(fb[j + 1:], fb[j])
            j += 1
        if len(picked) >= n_peaks:
            break
    if examined is not None:
        examined.append(visited)
    return np.asarray(order[picked], dtype=np.int64)


def peak_preserving_downsample(
    frequency: np.ndarray,
    power: np.ndarray,
    n_points: int,
    *,
    sense: Literal["max", "min"] = "max",
) -> tuple[np.ndarray, np.ndarray]:
    """``(frequency, power)`` thinned to ``n_points`` samples, as float32: the samples are
    cut into ``n_points`` runs of consecutive indices (run ``i`` starts at
    ``floor(i*n/n_points)``) and each run keeps its most significant sample, the first of
    equals (NaN counts as least significant). A spectrum of at most ``n_points`` samples
    is returned whole."""
    if sense not in ("max", "min"):
        raise ValueError(f"sense must be 'max' or 'min', not {sense!r}")
    n, m = int(frequency.size), int(n_points)
    if m <= 0:
        return np.empty(0, dtype=np.float32), np.empty(0, dtype=np.float32)
    if n <= m:
        return frequency.astype(np.float32), power.astype(np.float32)
    score = np.asarray(power if sense == "max" else -power, dtype=np.float64)
    score = np.where(np.isnan(score), -np.inf, score)
    first = (np.arange(m + 1, dtype=np.int64) * n) // m
    width = int(np.diff(first).max())
    # one row per run, padded on the right with -inf so a pad is never the first maximum
    cols = first[:-1, None] + np.arange(width, dtype=np.int64)[None, :]
    inside = cols < first[1:, None]
    table = np.where(inside, score[np.minimum(cols, n - 1)], -np.inf)
    picks = first[:-1] + np.argmax(table, axis=1)
    return frequency[picks].astype(np.float32), power[picks].astype(np.float32)


__all__ = [
    "SIDEREAL_FREQ_CPD",
    "SOLAR_FREQ_CPD",
    "YEARLY_FREQ_CPD",
    "is_alias_related",
    "local_maxima",
    "peak_preserving_downsample",
    "select_diverse_peaks",
    "select_top_peaks",
]
