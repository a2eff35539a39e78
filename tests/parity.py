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

import numpy as np

from oracle.ties import BoxInputs, prove_same_points  # noqa: F401 (re-exported)
from oracle.ties import box_members, fold_bins, prove_cross_period_tie, recover_boxes  # noqa: F401

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
