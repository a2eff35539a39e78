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
    at k = 1, 2, then the yearly sidelobe), as float64 columns."""
    orders = np.arange(2, harmonic_max + 1, dtype=np.float64)[:, None]
    day = np.array([SIDEREAL_FREQ_CPD, SOLAR_FREQ_CPD], dtype=np.float64)
    offsets = np.concatenate([(np.array([1.0, 2.0])[None, :] * day[:, None]).ravel(),
                              [YEARLY_FREQ_CPD]])[:, None]
    return orders, offsets


def _related(f_new, f_chosen: np.ndarray, tol: float, separation: float,
             orders: np.ndarray, offsets: np.ndarray) -> np.ndarray:
    """Element-wise over ``f_chosen``: does ``f_new`` duplicate, harmonise or alias it
    (the test of :func:`is_alias_related`, one chosen frequency per column)."""
    near = np.abs(f_new - f_chosen) < separation
    harmonic = np.abs(f_new - orders * f_chosen) < orders * tol
    subharmonic = np.abs(f_new - f_chosen / orders) < tol
    below = np.abs(f_new - np.abs(f_chosen - offsets)) < tol
    above = np.abs(f_new - (f_chosen + offsets)) < tol
    return near | (harmonic | subharmonic).any(0) | (below | above).any(0)


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
    return bool(_related(float(f_new), np.array([float(f_chosen)]), tol, sep, orders,
                         offsets)[0])


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
    candidates visited."""
    if n_peaks <= 0:
        return np.empty(0, dtype=np.int64)
    orders, offsets = _alias_lattice(harmonic_max)
    sep = tol if min_separation is None else min_separation
    picked = np.empty(0, dtype=np.int64)
    picked_f = np.empty(0, dtype=np.float64)
    visited = 0
    for idx in _descending(score, candidates):
        visited += 1
        f = float(frequency[idx])
        if picked.size and bool(
            (_related(f, picked_f, tol, sep, orders, offsets)
             | _related(picked_f, f, tol, sep, orders, offsets)).any()
        ):
            continue
        picked = np.append(picked, np.int64(idx))
        picked_f = np.append(picked_f, f)
        if picked.size >= n_peaks:
            break
    if examined is not None:
        examined.append(visited)
    return picked


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
