"""Brute-force BLS for tiny inputs -- TEST INFRASTRUCTURE ONLY.

An independent check of the oracle's prefix-sum sweep: every box statistic is an
*exactly rounded* direct sum (``math.fsum``) over the points the box holds, with no
cumulative sums at all. The binning follows the reference's definition
(src/cuperiod/methods/_bls_core.py:358-383): the fold ``w = x - P*floor(x/P)``,
``ind = int(w/bd) + 1`` clipped to ``[0, width-1]``, and the wrap COPY of
``oversample`` bins applied in order (a copy may read a bin an earlier copy wrote).
The objectives are the textbook weighted BLS statistics (Kovacs, Zucker & Mazeh 2002):
depth = mean(out) - mean(in), SNR = depth / sqrt(1/W_in + 1/W_out), log-likelihood
= W_in * depth^2 / 2; the search keeps the first strict maximum in (duration index,
start) order (``_bls_core.py:434``).

O(periods x durations x bins x N): use a few dozen points and periods.
"""

from __future__ import annotations

import math

import numpy as np

from oracle import OUTPUT_NAMES, ivar_floor


def _members(tau, period, bd, oversample, width):
    """Point indices held by each bin ``0..n_bins`` after the fold and the wrap copy."""
    n_bins = int(math.ceil(period / bd)) + oversample
    if n_bins > width - 1:
        n_bins = width - 1
    bins: list[list[int]] = [[] for _ in range(width)]
    for j, x in enumerate(tau.tolist()):
        w = x - period * math.floor(x / period)
        ind = int(w / bd) + 1
        ind = min(max(ind, 0), width - 1)
        bins[ind].append(j)
    for k in range(oversample):
        dst = n_bins - oversample + k
        if 0 <= dst < width:
            bins[dst] = list(bins[1 + k])
    return bins, n_bins


def bls_bruteforce(
    t, y, ivar, periods, *, bin_duration, dur_bins, oversample, width, obj_flag
) -> dict[str, np.ndarray]:
    """Per-period best box with exactly summed statistics (same outputs as the oracle)."""
    t = np.asarray(t, dtype=np.float64)
    yw = np.asarray(y, dtype=np.float64) * np.asarray(ivar, dtype=np.float64)
    wv = np.asarray(ivar, dtype=np.float64)
    t_min = float(t.min())
    tau = t - t_min
    sum_y = math.fsum(yw.tolist())
    sum_ivar = math.fsum(wv.tolist())
    floor = ivar_floor(int(t.size), sum_ivar)
    periods = np.asarray(periods, dtype=np.float64)
    out = {k: np.zeros(periods.size) for k in OUTPUT_NAMES}
    out["power"][:] = -np.inf
    ywl, wvl = yw.tolist(), wv.tolist()
    for p, period in enumerate(periods.tolist()):
        bins, n_bins = _members(tau, period, bin_duration, oversample, width)
        best = -math.inf
        best_nd = None
        for d in dur_bins:
            if d < 1 or d >= n_bins:
                continue
            for n in range(0, n_bins - d + 1):
                idx = [j for b in range(n + 1, n + d + 1) for j in bins[b]]
                y_in = math.fsum(ywl[j] for j in idx)
                iv_in = math.fsum(wvl[j] for j in idx)
                iv_out = sum_ivar - iv_in
                if iv_in < floor or iv_out < floor:
                    continue
                yin = y_in / iv_in
                yout = (sum_y - y_in) / iv_out
                if yout < yin:
                    continue
                depth = yout - yin
                if obj_flag == 0:
                    obj = depth / math.sqrt(1.0 / iv_in + 1.0 / iv_out)
                else:
                    obj = 0.5 * iv_in * depth * depth
                if obj > best:
                    best = obj
                    best_nd = (n, d, y_in, iv_in, iv_out)
        if best_nd is None:
            continue
        n, d, y_in, iv_in, iv_out = best_nd
        yin = y_in / iv_in
        yout = (sum_y - y_in) / iv_out
        depth = yout - yin
        derr = math.sqrt(1.0 / iv_in + 1.0 / iv_out)
        dsnr = depth / derr
        ll = 0.5 * iv_in * depth * depth
        dur = d * bin_duration
        out["power"][p] = dsnr if obj_flag == 0 else ll
        out["depth"][p] = depth
        out["depth_err"][p] = derr
        out["depth_snr"][p] = dsnr
        out["duration"][p] = dur
        out["transit_time"][p] = math.fmod(n * bin_duration + 0.5 * dur, period) + t_min
        out["log_likelihood"][p] = ll
    return out


def box_table(t, y, ivar, period, *, bin_duration, dur_bins, oversample, width):
    """Every box of one period: ``[(di, n, y_in, iv_in), ...]`` with exact sums.

    For tests that check individual box statistics (not just the maximum).
    """
    t = np.asarray(t, dtype=np.float64)
    tau = t - float(t.min())
    yw = (np.asarray(y, dtype=np.float64) * np.asarray(ivar, dtype=np.float64)).tolist()
    wv = np.asarray(ivar, dtype=np.float64).tolist()
    bins, n_bins = _members(tau, float(period), bin_duration, oversample, width)
    rows = []
    for di, d in enumerate(dur_bins):
        if d < 1 or d >= n_bins:
            continue
        for n in range(0, n_bins - d + 1):
            idx = [j for b in range(n + 1, n + d + 1) for j in bins[b]]
            rows.append(
                (di, n, math.fsum(yw[j] for j in idx), math.fsum(wv[j] for j in idx))
            )
    return rows, n_bins


def joint_bruteforce(bands, periods, *, bin_duration, dur_bins, oversample, width,
                     obj_flag):
    """The joint-ephemeris sweep with exactly summed box statistics.

    ``bands`` is an :class:`oracle.multiband.JointBands`. Every band is folded against
    the common origin; box ``(n, d)`` holds bins ``n+1 .. n+d`` of every band; the box
    value is the sum over the admissible, dipping bands of ``depth^2 / (1/W_in +
    1/W_out)`` (snr) or ``W_in depth^2 / 2`` (likelihood); a box counts when one band
    dips; first strict maximum in (duration index, start) order. Returns
    ``(best_n, best_d, found, y_in, ivar_in)`` like :func:`oracle.multiband.joint_kernel`.
    """
    periods = np.asarray(periods, dtype=np.float64)
    k_all = bands.n_bands
    n_p = int(periods.size)
    o_n = np.zeros(n_p, dtype=np.int64)
    o_d = np.zeros(n_p, dtype=np.int64)
    found = np.zeros(n_p, dtype=bool)
    o_y = np.zeros((k_all, n_p))
    o_i = np.zeros((k_all, n_p))
    per_band = []
    for k in range(k_all):
        sl = slice(int(bands.starts[k]), int(bands.starts[k + 1]))
        per_band.append((bands.tau[sl], bands.yw[sl].tolist(), bands.ivar[sl].tolist()))
    sums_y = [math.fsum(yw) for _, yw, _ in per_band]
    sums_i = [math.fsum(iv) for _, _, iv in per_band]
    for p, period in enumerate(periods.tolist()):
        folded = [_members(tau, period, bin_duration, oversample, width)
                  for tau, _, _ in per_band]
        n_bins = folded[0][1]
        best = -math.inf
        best_box = None
        for d in dur_bins:
            if d < 1 or d >= n_bins or d >= width:
                continue
            for n in range(0, n_bins - d + 1):
                total = 0.0
                hit = False
                sums = []
                for k in range(k_all):
                    bins = folded[k][0]
                    idx = [j for b in range(n + 1, n + d + 1) for j in bins[b]]
                    y_in = math.fsum(per_band[k][1][j] for j in idx)
                    iv_in = math.fsum(per_band[k][2][j] for j in idx)
                    sums.append((y_in, iv_in))
                    iv_out = sums_i[k] - iv_in
                    if iv_in < bands.floor[k] or iv_out < bands.floor[k]:
                        continue
                    yin = y_in / iv_in
                    yout = (sums_y[k] - y_in) / iv_out
                    if not yout >= yin:
                        continue
                    depth = yout - yin
                    hit = True
                    if obj_flag == 0:
                        total += depth * depth / (1.0 / iv_in + 1.0 / iv_out)
                    else:
                        total += 0.5 * iv_in * depth * depth
                if hit and total > best:
                    best = total
                    best_box = (n, d, sums)
        if best_box is None:
            continue
        n, d, sums = best_box
        found[p] = True
        o_n[p] = n
        o_d[p] = d
        for k, (y_in, iv_in) in enumerate(sums):
            o_y[k, p] = y_in
            o_i[k, p] = iv_in
    return o_n, o_d, found, o_y, o_i
