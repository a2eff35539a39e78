"""Comparison helpers for HIP-vs-oracle parity tests.

The bar (stated in every test that uses it):
* power: identical -inf pattern; every finite power within ``rtol`` (default 1e-6, the
  bound the engine is held to) of the FP64 oracle -- observed ~1e-13, since the engine's
  bins are exact and the oracle's carry FP64 rounding;
* the chosen box (duration, transit time) identical, or PROVEN to hold the same
  multiset of points as the oracle's box, so that their exact statistics are equal and
  the choice is a tie: both boxes are recovered from the outputs and their members
  listed by folding every point with the oracle's formula, wrap copy included
  (:func:`box_members`); callers without ``inputs`` need identical boxes everywhere;
* the best period identical: the same argmax index, and all seven outputs bit-identical
  there (the engine searches each segment's near-maximum periods again with the
  reference's FP64 sequence);
* powers below 1e-290 (underflowed objectives) are compared in value only.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

KEYS = ("power", "depth", "depth_err", "depth_snr", "duration", "transit_time",
        "log_likelihood")


def rel_err(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both = np.isfinite(a) & np.isfinite(b)
    out = np.zeros(a.shape)
    scale = np.maximum(np.abs(b[both]), 1e-300)
    out[both] = np.abs(a[both] - b[both]) / scale
    return out


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
        import oracle

        t = np.asarray(t, dtype=np.float64)
        t_min = float(t.min())
        per, bd, wd, db = [], [], [], []
        for periods, durations in specs:
            periods = np.asarray(periods, dtype=np.float64)
            if periods.size == 0:
                continue
            p, b, dbins, width = oracle.search_spec(periods, durations, oversample)
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
                           name: str = "") -> None:
    """Assert that periods ``i`` and ``j`` tie exactly in ``ref`` (the oracle's own
    powers are bit-identical) and that their boxes hold the same multiset of points, so
    their exact statistics are equal: the reference's choice between them is decided
    by numpy's (unstable) sort order of equal scores."""
    assert ref["power"][i] == ref["power"][j], (
        f"{name}: periods {i} and {j}: powers {ref['power'][i]!r} vs {ref['power'][j]!r}")
    bi, bj = recover_boxes(ref, inp, i), recover_boxes(ref, inp, j)
    assert bi and bj, f"{name}: boxes not recoverable at {i}, {j}"
    mi = [box_members(inp, i, n, d) for n, d in bi]
    mj = [box_members(inp, j, n, d) for n, d in bj]
    assert any(np.array_equal(x, y) for x in mi for y in mj), (
        f"{name}: periods {i} and {j} have equal powers but boxes of different points")


def compare(got: dict, ref: dict, *, inputs: BoxInputs | None = None, rtol: float = 1e-6,
            box_frac: float = 0.995, name: str = "", best_identical: bool = True) -> dict:
    """Assert the parity bar; return the measured errors (and the number of periods
    whose differing boxes were proven exact ties)."""
    gp = np.asarray(got["power"])
    rp = np.asarray(ref["power"])
    assert gp.shape == rp.shape, name
    assert np.array_equal(np.isneginf(gp), np.isneginf(rp)), (
        f"{name}: -inf pattern differs at {np.flatnonzero(np.isneginf(gp) != np.isneginf(rp))[:10]}"
    )
    assert np.array_equal(np.isfinite(gp), np.isfinite(rp)), name
    perr = rel_err(gp, rp)
    worst = int(np.argmax(perr)) if perr.size else 0
    assert perr.size == 0 or perr.max() <= rtol, (
        f"{name}: power rel err {perr.max():.3g} at {worst}: {gp[worst]!r} vs {rp[worst]!r}"
    )
    same_box = ((np.asarray(got["duration"]) == np.asarray(ref["duration"]))
                & (np.asarray(got["transit_time"]) == np.asarray(ref["transit_time"])))
    # Powers that underflowed (e.g. a log-likelihood of flux scaled by 1e-200) tie at
    # zero or subnormal values, where either order is right; they are left out of the
    # box-identity fraction (their power must still agree).
    meaningful = np.isfinite(rp) & (np.abs(rp) >= 1e-290)
    frac = float(same_box[meaningful].mean()) if meaningful.any() else 1.0
    diff = np.flatnonzero(~same_box & meaningful)
    proven = 0
    if inputs is None:
        assert frac >= box_frac, f"{name}: same box at only {frac:.4f} of periods"
        assert diff.size == 0, f"{name}: {diff.size} differing boxes and no inputs to prove ties"
    elif diff.size:
        # every differing box is proven to hold the oracle's box's points: the fraction
        # of identical boxes is then a measurement, not a bar
        proven = prove_same_points(got, ref, inputs, diff, name)
    underflowed = not (np.abs(rp[np.isfinite(rp)]) >= 1e-290).any() if gp.size else True
    if best_identical and gp.size and np.isfinite(rp).any() and not underflowed:
        gi = int(np.argmax(np.where(np.isfinite(gp), gp, -np.inf)))
        ri = int(np.argmax(np.where(np.isfinite(rp), rp, -np.inf)))
        assert gi == ri, f"{name}: best period index {gi} vs {ri}"
        for k in KEYS:
            assert np.asarray(got[k])[gi] == np.asarray(ref[k])[ri], (
                f"{name}: {k} at the best period: {got[k][gi]!r} vs {ref[k][ri]!r}")
    errs = {"power": float(perr.max()) if perr.size else 0.0, "same_box": frac,
            "ties_proven": proven}
    for k in ("depth", "depth_snr", "log_likelihood"):
        e = rel_err(got[k], ref[k])[same_box]
        errs[k] = float(e.max()) if e.size else 0.0
    return errs
