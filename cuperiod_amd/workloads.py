"""Seeded synthetic survey workloads used by the benchmark and the tests.

Eight workloads, each generated with ``np.random.default_rng(20261203 + k)`` for the k-th
of the list below:

1. ``small_curve``: one curve, N = 2,000 over 100 d, box transit P = 3.7 d, depth 2e-3, duration 0.1 d,
   phase 0.3; 50,000 frequencies linear in [0.02, 2] /d; durations {0.05, 0.1, 0.15,
   0.2} d; oversample 10;
2. ``gapped_curve``: one curve, N = 20,000 over 1,000 d with 30% removed in seasonal gaps,
   sigma 2e-3, P = 11.3 d, depth 1.5e-3, duration 0.2 d; 1,000,000 frequencies in
   [0.001, 5] /d; 8 durations log-spaced in [0.04, 0.4] d;
3. ``survey_batch``: 4,096 curves, N log-uniform in [100, 5,000], baselines uniform in [50, 2,000] d,
   sigma log-uniform in [5e-4, 5e-3]; even-index curves carry a box transit (P
   log-uniform in [0.5, 20] d, depth log-uniform in [5e-4, 1e-2], duration/period
   uniform in [0.01, 0.05]); 200,000 frequencies in [1/baseline, 2] /d per curve;
   6 period-scaled durations per log segment;
4. ``learned_batch``: 16,384 curves of the gapped-curve recipe (2.), each with its own
   transit period (log-uniform in [2, 20] d) and phase, depth 1.5e-3 and duration 0.2 d;
   200,000 frequencies in [1/1,000, 2] /d; 6 period-scaled durations per log segment
   (the learned proposer's workload; its training curves use other seeds);
5. ``long_baseline_batch``: 256 curves, N = 100,000 over 10,000 d offset by 2.45e6 d, P = 0.7 d, depth 1e-3,
   duration 0.05 d; 4,000,000 frequencies in [1e-4, 10] /d; 16 durations;
6. ``multiband_survey``: stars of 6 bands x 150 points each over 1,095 d (random
   epochs offset by 2.46e6 d), band noise 2-8 mmag-like (sigma 2e-3 .. 8e-3 in flux),
   chromatic box transits (P log-uniform in [0.5, 20] d, depth 0.7-1.3 x a
   log-uniform 6e-3 .. 2e-2, duration 0.02-0.05 of P) in every other star; searched
   jointly over the default BLS grid of a 1,095 d baseline;
8. ``published_curve`` (seed k = 8): the shape of cuPeriod's published single-curve BLS
   timing (benchmarks/benchmark.py:36-47, 149-156): N = 891 over a 2,175 d baseline,
   sigma 1e-3, a 1.3 d box transit of depth 3e-3 and duration 0.08 d, periods 0.5-4 d on
   the default grid with grid_oversample 1 (76,125 periods here, 76,124 published);
7. ``tls_survey``: TESS-sector-like curves, 20,000 points over 27 d (offset 2.459e6 d),
   sigma log-uniform in [2e-4, 2e-3], limb-darkened transits (u1 0.4, u2 0.3; P
   log-uniform in [0.6, 12] d, depth log-uniform in [3e-4, 5e-3], duration 0.02-0.06 of
   P) in every other curve; the default TLS grid of a 27 d baseline.

A curve is ``(t, flux, dy)`` with flux in the flux domain (a transit is a dip) and a
constant ``dy`` per curve (uniform weights).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SEED = 20261203


def _box_dip(t: np.ndarray, period: float, t0: float, duration: float) -> np.ndarray:
    """Points inside a box of ``duration`` centred on ``t0 + k*period``."""
    ph = np.mod((t - t0) / period + 0.5, 1.0) - 0.5
    return np.abs(ph) < 0.5 * duration / period


@dataclass
class Workload:
    """Light curves plus how to search them."""

    name: str
    curves: list
    grids: list
    n_durations: int | None = None
    durations: np.ndarray | None = None
    oversample: int = 10
    truth: list = field(default_factory=list)


def small_curve() -> Workload:
    rng = np.random.default_rng(SEED + 1)
    n = 2000
    t = np.sort(rng.uniform(0.0, 100.0, n))
    y = 1.0 + rng.normal(0.0, 1e-3, n)
    y[_box_dip(t, 3.7, 0.3 * 3.7, 0.1)] -= 2e-3
    from cuperiod_amd.batch.engine import LinspaceGrid

    return Workload("small_curve", [(t, y, np.full(n, 1e-3))], [LinspaceGrid(0.02, 2.0, 50000)],
                    durations=np.array([0.05, 0.1, 0.15, 0.2]), truth=[3.7])


def gapped_curve() -> Workload:
    rng = np.random.default_rng(SEED + 2)
    t = np.sort(rng.uniform(0.0, 1000.0, 20000))
    t = t[np.mod(t / 365.25, 1.0) >= 0.3]
    n = t.size
    y = 1.0 + rng.normal(0.0, 2e-3, n)
    y[_box_dip(t, 11.3, 0.0, 0.2)] -= 1.5e-3
    from cuperiod_amd.batch.engine import LinspaceGrid

    return Workload("gapped_curve", [(t, y, np.full(n, 2e-3))], [LinspaceGrid(0.001, 5.0, 1_000_000)],
                    durations=np.geomspace(0.04, 0.4, 8), truth=[11.3])


def survey_curves(n_curves: int, seed: int, n_freq: int = 200_000):
    """Uneven survey curves with per-curve linspace grids (survey and proposer batches)."""
    from cuperiod_amd.batch.engine import LinspaceGrid

    rng = np.random.default_rng(seed)
    n_pts = np.round(np.exp(rng.uniform(np.log(100), np.log(5000), n_curves))).astype(int)
    base = rng.uniform(50.0, 2000.0, n_curves)
    sig = np.exp(rng.uniform(np.log(5e-4), np.log(5e-3), n_curves))
    curves, grids, truth = [], [], []
    for c in range(n_curves):
        t = np.sort(rng.uniform(0.0, base[c], n_pts[c]))
        y = 1.0 + rng.normal(0.0, sig[c], n_pts[c])
        p = None
        if c % 2 == 0:
            p = float(np.exp(rng.uniform(np.log(0.5), np.log(20.0))))
            depth = float(np.exp(rng.uniform(np.log(5e-4), np.log(1e-2))))
            q = float(rng.uniform(0.01, 0.05))
            t0 = float(rng.uniform(0.0, p))
            y[_box_dip(t, p, t0, q * p)] -= depth
        curves.append((t, y, np.full(n_pts[c], sig[c])))
        grids.append(LinspaceGrid(1.0 / base[c], 2.0, n_freq))
        truth.append(p)
    return curves, grids, truth


def survey_batch(n_curves: int = 4096, n_freq: int = 200_000) -> Workload:
    curves, grids, truth = survey_curves(n_curves, SEED + 3, n_freq)
    return Workload("survey_batch", curves, grids, n_durations=6, truth=truth)


def gapped_curves(n_curves: int, seed: int, n_freq: int = 200_000):
    """``(curves, grids, periods)``: gapped-curve recipe curves, N = 20,000 uniform over
    1,000 d with the first 30% of every 365.25 d season removed, sigma 2e-3, a 0.2 d box
    transit of depth 1.5e-3 at a period log-uniform in [2, 20] d and a uniform phase;
    ``np.linspace(1/1000, 2, n_freq)`` frequencies each."""
    from cuperiod_amd.batch.engine import LinspaceGrid

    rng = np.random.default_rng(seed)
    curves, grids, periods = [], [], []
    for _ in range(n_curves):
        t = np.sort(rng.uniform(0.0, 1000.0, 20_000))
        t = t[np.mod(t / 365.25, 1.0) >= 0.3]
        y = 1.0 + rng.normal(0.0, 2e-3, t.size)
        p = float(np.exp(rng.uniform(np.log(2.0), np.log(20.0))))
        y[_box_dip(t, p, float(rng.uniform(0.0, p)), 0.2)] -= 1.5e-3
        curves.append((t, y, np.full(t.size, 2e-3)))
        grids.append(LinspaceGrid(1.0 / 1000.0, 2.0, n_freq))
        periods.append(p)
    return curves, grids, periods


def learned_batch(n_curves: int = 16384, n_freq: int = 200_000) -> Workload:
    curves, grids, truth = gapped_curves(n_curves, SEED + 4, n_freq)
    return Workload("learned_batch", curves, grids, n_durations=6, truth=truth)


def long_baseline_batch(n_curves: int = 256, n_freq: int = 4_000_000, n_points: int = 100_000) -> Workload:
    from cuperiod_amd.batch.engine import LinspaceGrid

    rng = np.random.default_rng(SEED + 5)
    curves, grids = [], []
    for _ in range(n_curves):
        t = np.sort(rng.uniform(0.0, 1e4, n_points)) + 2.45e6
        y = 1.0 + rng.normal(0.0, 1e-3, n_points)
        y[_box_dip(t, 0.7, 2.45e6 + 0.1, 0.05)] -= 1e-3
        curves.append((t, y, np.full(n_points, 1e-3)))
        grids.append(LinspaceGrid(1e-4, 10.0, n_freq))
    return Workload("long_baseline_batch", curves, grids, n_durations=16, truth=[0.7] * n_curves)




def multiband_survey(n_stars: int = 64, n_bands: int = 6, n_per_band: int = 150,
                     span: float = 1095.0):
    """``(stars, truth)``: ``stars[i]`` is a list of ``(t, flux, dy)`` per band."""
    rng = np.random.default_rng(SEED + 6)
    stars, truth = [], []
    for i in range(n_stars):
        bands = []
        period = float(np.exp(rng.uniform(np.log(0.5), np.log(20.0))))
        # deep enough for the six-band stack to rise above the noise maximum of ~2e5
        # periods with 3-8 in-transit points per 150-point band
        depth = float(np.exp(rng.uniform(np.log(6e-3), np.log(2e-2))))
        frac = float(rng.uniform(0.02, 0.05))
        t0 = 2.46e6 + float(rng.uniform(0.0, period))
        for b in range(n_bands):
            t = 2.46e6 + np.sort(rng.uniform(0.0, span, n_per_band))
            sigma = float(rng.uniform(2e-3, 8e-3))
            y = 1.0 + rng.normal(0.0, sigma, n_per_band)
            if i % 2 == 0:
                scale = 0.7 + 0.6 * b / max(1, n_bands - 1)
                y[_box_dip(t, period, t0, frac * period)] -= depth * scale
            bands.append((t, y, np.full(n_per_band, sigma)))
        stars.append(bands)
        truth.append(period if i % 2 == 0 else None)
    return stars, truth


def tls_survey(n_curves: int = 256, n_points: int = 20_000, span: float = 27.0):
    """``(curves, truth)``: ``curves[i] = (t, flux, dy)``."""
    rng = np.random.default_rng(SEED + 7)
    curves, truth = [], []
    for i in range(n_curves):
        t = 2.459e6 + np.sort(rng.uniform(0.0, span, n_points))
        sigma = float(np.exp(rng.uniform(np.log(2e-4), np.log(2e-3))))
        y = 1.0 + rng.normal(0.0, sigma, n_points)
        period = None
        if i % 2 == 0:
            period = float(np.exp(rng.uniform(np.log(0.6), np.log(12.0))))
            depth = float(np.exp(rng.uniform(np.log(3e-4), np.log(5e-3))))
            q = float(rng.uniform(0.02, 0.06))
            t0 = 2.459e6 + float(rng.uniform(0.0, period))
            ph = ((t - t0) / period) % 1.0
            x = np.where(ph > 0.5, ph - 1.0, ph)
            z = np.clip(2.0 * np.abs(x) / q, 0.0, 1.0)
            mu = np.sqrt(1.0 - z * z)
            prof = 1.0 - 0.4 * (1.0 - mu) - 0.3 * (1.0 - mu) ** 2
            inside = np.abs(x) < q / 2.0
            y[inside] -= depth * prof[inside] / prof.max()
        curves.append((t, y, np.full(n_points, sigma)))
        truth.append(period)
    return curves, truth





def published_curve():
    """``(t, flux, dy, settings)`` of the published single-curve BLS timing's shape."""
    from cuperiod_amd.core.config import BLSSettings

    rng = np.random.default_rng(SEED + 8)
    t = np.sort(rng.uniform(0.0, 2175.0, 891))
    t[0], t[-1] = 0.0, 2175.0
    y = 1.0 + rng.normal(0.0, 1e-3, t.size)
    y[_box_dip(t, 1.3, 0.4, 0.08)] -= 3e-3
    settings = BLSSettings(min_period_days=0.5, max_period_days=4.0, grid_oversample=1)
    return t, y, np.full(t.size, 1e-3), settings


__all__ = ["SEED", "Workload", "gapped_curve", "long_baseline_batch", "multiband_survey",
           "published_curve", "gapped_curves", "learned_batch",
           "small_curve", "survey_batch", "survey_curves", "tls_survey"]
