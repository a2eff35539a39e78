"""Exact-tie proofs between two BLS results -- TEST INFRASTRUCTURE, never the product.

Used by ``tests/`` and ``bench.py``'s CPU-baseline leg, as the checker. Two searches of
the same light curve may pick different boxes at a period only when the boxes hold the
same points: their exact statistics are then equal, and which one wins depends on the
rounding of the sums (the reference's sequential FP64 prefix sums, the engine's exact
integer sums). :func:`prove_same_points` recovers both boxes from the outputs, lists
their members by folding every point with the reference's formula (the wrap copy
included, src/cuperiod/methods/_bls_core.py:369-390) and compares the multisets.

Pinned by ``tests/test_oracle.py``: the members of the oracle's own box reproduce its
power by direct summation; a box moved by one bin onto a different point is refused.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class BoxInputs:
    """What a search was run on, per period: enough to list the points of any box.

    ``tau`` = times minus ``t_min`` (float64, as the oracle forms them); ``period``,
    ``bin_duration``, ``width`` and ``dur_bins`` (a list of int arrays, one per period,
    or one shared array) describe each period's binning (``_search_spec``)."""

    tau: np.ndarray
    t_min: float
    period: np.ndarray
    bin_duration: np.ndarray
    width: np.ndarray
    oversample: int
    dur_bins: list

    @classmethod
    def for_specs(cls, t, specs, oversample: int) -> "BoxInputs":
        """From absolute times and ``[(periods, durations), ...]`` search specs."""
        from oracle import search_spec

        t = np.asarray(t, dtype=np.float64)
        t_min = float(t.min())
        per, bd, wd, db = [], [], [], []
        for periods, durations in specs:
            periods = np.asarray(periods, dtype=np.float64)
            if periods.size == 0:
                continue
            p, b, dbins, width = search_spec(periods, durations, oversample)
            per.append(p)
            bd.append(np.full(p.size, b))
            wd.append(np.full(p.size, width, dtype=np.int64))
            db.extend([np.asarray(dbins, dtype=np.int64)] * p.size)
        return cls(tau=t - t_min, t_min=t_min, period=np.concatenate(per),
                   bin_duration=np.concatenate(bd), width=np.concatenate(wd),
                   oversample=int(oversample), dur_bins=db)

    def n_bins(self, i: int) -> int:
        P, bd = float(self.period[i]), float(self.bin_duration[i])
        return min(int(math.ceil(P / bd)) + self.oversample, int(self.width[i]) - 1)


def fold_bins(inp: BoxInputs, i: int) -> np.ndarray:
    """Each point's phase bin at period ``i`` by the oracle's formula:
    ``int((x - P*floor(x/P))/bd) + 1`` clipped to ``[0, width-1]``
    (src/cuperiod/methods/_bls_core.py:369-378)."""
    P, bd = float(inp.period[i]), float(inp.bin_duration[i])
    w = inp.tau - P * np.floor(inp.tau / P)
    return np.clip((w / bd).astype(np.int64) + 1, 0, int(inp.width[i]) - 1)


def box_members(inp: BoxInputs, i: int, n: int, d: int,
                bins: np.ndarray | None = None) -> np.ndarray:
    """Sorted indices (with multiplicity) of the points in box (start ``n``, ``d`` bins)
    of period ``i``: the prefix difference over bins ``n+1 .. n+d`` after the wrap copy
    (src/cuperiod/methods/_bls_core.py:379-390), each bin standing for the original bin
    whose points it holds."""
    width = int(inp.width[i])
    nb = inp.n_bins(i)
    ind = fold_bins(inp, i) if bins is None else bins
    src = np.arange(nb + 1)
    for k in range(inp.oversample):  # in order: a copy may read a bin copied before
        dst = nb - inp.oversample + k
        if 0 <= dst < width:
            src[dst] = src[1 + k]
    mult = np.bincount(src[n + 1:n + d + 1], minlength=width)
    return np.repeat(np.arange(ind.size), mult[ind])


def recover_boxes(res: dict, inp: BoxInputs, i: int) -> list[tuple[int, int]]:
    """The ``(n, d)`` boxes of period ``i`` whose duration and transit time the result
    holds bit for bit (``duration = d*bd``, ``transit_time = fmod(n*bd + dur/2, P) +
    t_min``)."""
    P, bd = float(inp.period[i]), float(inp.bin_duration[i])
    dur = float(res["duration"][i])
    tt = float(res["transit_time"][i])
    nb = inp.n_bins(i)
    d = int(round(dur / bd))
    if d < 1 or float(d) * bd != dur:
        return []
    x = (tt - inp.t_min - 0.5 * dur) / bd
    found = []
    for base in (x, x + P / bd, x - P / bd):
        for n in range(int(math.floor(base)) - 2, int(math.floor(base)) + 3):
            if 0 <= n <= nb - d and (n, d) not in found:
                if math.fmod(float(n) * bd + 0.5 * dur, P) + inp.t_min == tt:
                    found.append((n, d))
    return found


def prove_same_points(got: dict, ref: dict, inp: BoxInputs, idx: np.ndarray,
                      name: str = "") -> int:
    """Assert that at each period of ``idx`` the two results' boxes hold the same
    multiset of points; returns how many periods were proven."""
    for i in np.asarray(idx, dtype=np.int64).tolist():
        a = recover_boxes(got, inp, i)
        b = recover_boxes(ref, inp, i)
        assert a and b, f"{name}: period {i}: box not recoverable ({a}, {b})"
        bins = fold_bins(inp, i)
        ma = [box_members(inp, i, n, d, bins) for n, d in a]
        mb = [box_members(inp, i, n, d, bins) for n, d in b]
        ok = any(np.array_equal(x, y) for x in ma for y in mb)
        assert ok, (f"{name}: period {i} (P={inp.period[i]!r}): boxes {a} and {b} hold "
                    "different points (not an exact tie)")
    return int(np.asarray(idx).size)


def prove_cross_period_tie(ref: dict, inp: BoxInputs, i: int, j: int,
                           name: str = "") -> int:
    """Assert that periods ``i`` and ``j`` tie in exact arithmetic: their boxes hold the
    same multiset of points, so their exact statistics are equal, and the oracle's own
    FP64 powers there differ by no more than the rounding of summing the same terms in
    another order (at most 1e-12 relative; the reference's choice between such periods
    is decided by that rounding and by numpy's unstable sort of equal scores). Returns
    the distance between the two oracle powers in units in the last place."""
    pi, pj = float(ref["power"][i]), float(ref["power"][j])
    ulps = int(round(abs(pi - pj) / np.spacing(max(abs(pi), abs(pj))))) if pi != pj else 0
    assert abs(pi - pj) <= 1e-12 * max(abs(pi), abs(pj)), (
        f"{name}: periods {i} and {j}: powers {pi!r} vs {pj!r} ({ulps} ulp apart)")
    bi, bj = recover_boxes(ref, inp, i), recover_boxes(ref, inp, j)
    assert bi and bj, f"{name}: boxes not recoverable at {i}, {j}"
    mi = [box_members(inp, i, n, d) for n, d in bi]
    mj = [box_members(inp, j, n, d) for n, d in bj]
    assert any(np.array_equal(x, y) for x in mi for y in mj), (
        f"{name}: periods {i} and {j} have powers {ulps} ulp apart but boxes of "
        "different points")
    return ulps


__all__ = ["BoxInputs", "box_members", "fold_bins", "prove_cross_period_tie",
           "prove_same_points", "recover_boxes"]
