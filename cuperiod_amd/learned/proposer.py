"""Learned candidate proposer plus fine refinement, with the plain search as fallback.

For each light curve on a linspace frequency grid (the plain, full-grid search):

1. a coarse BLS spectrum at 4,096 log-spaced frequencies spanning the grid (the same
   engine, period segments and durations as the plain search);
2. features: the coarse power standardised per curve and a 512-bin cadence histogram
   (``cpp_features``);
3. the network scores every coarse cell; greedy windows are taken around the most
   probable cells (``cpp_head_windows``) and kept, most probable first, until their
   probability mass reaches ``window_mass`` or ``max_windows`` are kept;
4. a curve whose kept mass is below ``threshold`` falls back to the full grid; every
   other curve is searched only at the full grid's frequencies within
   ``fine_per_window`` grid steps around each kept window (2,000 by default).

Refinement evaluates the plain grid's own frequencies with the plain search's segment
binning, so a refined curve's best period is the plain best period whenever a window
covers it. Refined and fallback curves share one engine launch.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Sequence

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.batch import engine
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.learned.network import L, ProposerNet, load_parameters
from cuperiod_amd.methods.bls import _log_segments

#: Weights trained on synthetic survey curves (see ``cuperiod_amd.learned.train``).
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "weights" / "proposer_survey.npz"
#: Coarse standardised power is clipped to +-FEATURE_CLIP.
FEATURE_CLIP = 30.0


def coarse_frequencies(f_lo: float, f_hi: float) -> np.ndarray:
    """The L log-spaced coarse frequencies of a grid spanning ``[f_lo, f_hi]``."""
    return np.exp(np.linspace(math.log(f_lo), math.log(f_hi), L))


def coarse_plan(c: int, grid: engine.LinspaceGrid, settings: BLSSettings) -> engine.CurvePlan:
    """Explicit-period plan of the coarse grid, split into the log period segments
    (and durations) the plain search uses on ``grid``."""
    periods = np.ascontiguousarray(1.0 / coarse_frequencies(grid.start, grid.stop)[::-1])
    plan = engine.CurvePlan()
    for per, durations in _log_segments(periods, settings):
        plan.segments.append(
            engine._SegPlan(curve=c, n_periods=int(per.size), durations=durations,
                            p_max=float(per.max()), mode=0, periods=np.ascontiguousarray(per)))
        plan.n_periods += int(per.size)
    return plan


def coarse_position(freq: np.ndarray, f_lo: np.ndarray, f_hi: np.ndarray) -> np.ndarray:
    """Nearest coarse cell (0..L-1, ascending frequency) of each frequency."""
    step = (np.log(f_hi) - np.log(f_lo)) / (L - 1)
    x = np.rint((np.log(freq) - np.log(f_lo)) / step)
    return np.clip(x, 0, L - 1).astype(np.int64)


def window_geometry(grids, fine_per_window: int):
    """Per-curve ``(f_lo, dlog, half_width)``: the coarse cells' log spacing and the
    frequency half-width of a refinement window of ``fine_per_window`` grid steps."""
    f_lo = np.array([g.start for g in grids], dtype=np.float64)
    f_hi = np.array([g.stop for g in grids], dtype=np.float64)
    num = np.array([g.num for g in grids], dtype=np.float64)
    dlog = (np.log(f_hi) - np.log(f_lo)) / (L - 1)
    half = 0.5 * fine_per_window * (f_hi - f_lo) / (num - 1)
    return f_lo, dlog, half


def label_ranges(freq: np.ndarray, f_lo: np.ndarray, dlog: np.ndarray,
                 half: np.ndarray) -> np.ndarray:
    """``[n, 2]`` cell ranges whose frequencies lie within ``half`` of ``freq`` (at least
    the nearest cell)."""
    lo_f = np.maximum(freq - half, f_lo)
    lo = np.ceil(np.log(lo_f / f_lo) / dlog - 1e-9)
    hi = np.floor(np.log((freq + half) / f_lo) / dlog + 1e-9)
    near = np.rint(np.log(freq / f_lo) / dlog)
    lo = np.minimum(lo, near)
    hi = np.maximum(hi, near)
    out = np.stack([np.clip(lo, 0, L - 1), np.clip(hi, 0, L - 1)], axis=1)
    return out.astype(np.int32)


def features(pb: engine.PackedBatch, out, summary):
    """``[n, L, 2]`` FP32 device features of a coarse search's outputs."""
    import torch

    cs = pb.curve_set
    device = out.device
    n = len(pb.plans)
    x = torch.empty((n, L, 2), dtype=torch.float32, device=device)
    if np.any(pb.curve_count != L):
        raise ValueError("features need exactly L coarse periods per curve")
    pbase = torch.from_numpy(cs.point_base).to(device)
    npts = torch.from_numpy(cs.n_points.astype(np.int32)).to(device)
    base = torch.from_numpy(cs.baseline).to(device)
    import ctypes

    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    lib = _lib.load_library()
    st = lib.cpp_features(stream, ctypes.c_void_p(out.data_ptr()),
                          ctypes.c_void_p(pb.curve_base_d.data_ptr()),
                          ctypes.c_void_p(summary.data_ptr()), ctypes.c_void_p(cs.tau.data_ptr()),
                          ctypes.c_void_p(pbase.data_ptr()), ctypes.c_void_p(npts.data_ptr()),
                          ctypes.c_void_p(base.data_ptr()), n, FEATURE_CLIP,
                          ctypes.c_void_p(x.data_ptr()))
    if st != _lib.CPB_OK:
        raise _lib.status_error(st, f"proposer features: status {st}")
    return x


def window_ranges(centres: np.ndarray, grid: engine.LinspaceGrid, f_lo: float, f_hi: float,
                  half: int) -> list[tuple[int, int]]:
    """Merged ascending-period index ranges of the full grid within ``half`` grid steps
    of each coarse window centre."""
    if centres.size == 0:
        return []
    freq = coarse_frequencies(f_lo, f_hi)[centres]
    step = (grid.stop - grid.start) / (grid.num - 1)
    i = np.rint((freq - grid.start) / step).astype(np.int64)
    k = grid.num - 1 - i  # ascending-period index
    lo = np.clip(k - half, 0, grid.num)
    hi = np.clip(k + half, 0, grid.num)
    order = np.argsort(lo)
    ranges: list[tuple[int, int]] = []
    for a, b in zip(lo[order], hi[order]):
        if b <= a:
            continue
        if ranges and a <= ranges[-1][1]:
            ranges[-1] = (ranges[-1][0], max(ranges[-1][1], int(b)))
        else:
            ranges.append((int(a), int(b)))
    return ranges


@dataclass
class LearnedResult:
    """Per-curve outcome of :meth:`LearnedProposer.search`."""

    best_period: np.ndarray
    best_power: np.ndarray
    fallback: np.ndarray
    confidence: np.ndarray
    n_fine: np.ndarray
    timings: dict[str, float] = field(default_factory=dict)


@dataclass
class PreparedGrids:
    """What depends only on the curves and grids: full and coarse plans, the coarse
    search's packed descriptors."""

    cs: engine.CurveSet
    grids: list
    full: list
    coarse: engine.PackedBatch
    geometry: tuple = ()


def prepare_grids(cs: engine.CurveSet, grids: Sequence[engine.LinspaceGrid],
                  settings: BLSSettings, *, full: list | None = None,
                  fine_per_window: int = 2000) -> PreparedGrids:
    """Full-grid plans (built here unless given), the coarse search's packed
    descriptors and the window geometry of a batch."""
    import torch

    grids = list(grids)
    if full is None:
        full = [engine.plan_grid(cs, c, g, settings) for c, g in enumerate(grids)]
    coarse = engine.pack_plans(cs, [coarse_plan(c, g, settings) for c, g in enumerate(grids)],
                               settings)
    dev = cs.tau.device
    geometry = tuple(torch.from_numpy(a).to(dev) for a in window_geometry(grids, fine_per_window))
    return PreparedGrids(cs=cs, grids=grids, full=full, coarse=coarse, geometry=geometry)


class LearnedProposer:
    """Coarse search, network windows, refinement and fallback over a batch."""

    def __init__(self, params: dict[str, np.ndarray] | None = None, *, max_windows: int = 32,
                 radius: int = 1, fine_per_window: int = 2000, window_mass: float = 0.995,
                 threshold: float = 0.9) -> None:
        if params is None:
            params = load_parameters(DEFAULT_WEIGHTS)
        self.net = ProposerNet(params)
        self.max_windows = int(max_windows)
        self.radius = int(radius)
        self.fine_per_window = int(fine_per_window)
        self.window_mass = float(window_mass)
        self.threshold = float(threshold)

    def prepare(self, cs: engine.CurveSet, grids: Sequence[engine.LinspaceGrid],
                settings: BLSSettings, full: list | None = None) -> PreparedGrids:
        """Plans that do not depend on the network's output (built once per batch)."""
        return prepare_grids(cs, grids, settings, full=full,
                             fine_per_window=self.fine_per_window)

    def propose(self, prep: PreparedGrids):
        """Coarse search, features and windows: ``(centre, mass)`` host arrays."""
        out, summary = engine.run_packed(prep.coarse)
        x = features(prep.coarse, out, summary)
        centre, mass = self.net.windows(x, self.max_windows, self.radius, *prep.geometry)
        return centre.cpu().numpy(), mass.cpu().numpy()

    def select(self, centre: np.ndarray, mass: np.ndarray):
        """Windows kept per curve (most probable first) and each curve's kept mass."""
        kept: list[np.ndarray] = []
        conf = np.zeros(centre.shape[0])
        for c in range(centre.shape[0]):
            ok = centre[c] >= 0
            cen, m = centre[c][ok], mass[c][ok]
            order = np.argsort(-m, kind="stable")
            cum = np.cumsum(m[order])
            n_keep = int(np.searchsorted(cum, self.window_mass) + 1) if cum.size else 0
            n_keep = min(n_keep, cum.size)
            kept.append(cen[order[:n_keep]])
            conf[c] = float(cum[n_keep - 1]) if n_keep else 0.0
        return kept, conf

    def search(self, prep: PreparedGrids, settings: BLSSettings) -> LearnedResult:
        """Learned proposal + refinement, fallback below ``threshold``."""
        import torch

        t0 = time.perf_counter()
        centre, mass = self.propose(prep)
        t1 = time.perf_counter()
        kept, conf = self.select(centre, mass)
        fallback = conf < self.threshold
        plans = []
        for c, grid in enumerate(prep.grids):
            if fallback[c]:
                plans.append(prep.full[c])
                continue
            ranges = window_ranges(kept[c], grid, grid.start, grid.stop, self.fine_per_window // 2)
            plans.append(engine.window_plan(prep.full[c], ranges))
        pb = engine.pack_plans(prep.cs, plans, settings)
        t2 = time.perf_counter()
        out, summary = engine.run_packed(pb)
        res = engine.finish(pb, out, summary)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        return LearnedResult(
            best_period=res.best_period, best_power=res.best_power, fallback=fallback,
            confidence=conf, n_fine=pb.curve_count.copy(),
            timings={"propose": t1 - t0, "plan": t2 - t1, "refine": t3 - t2},
        )


__all__ = ["DEFAULT_WEIGHTS", "LearnedProposer", "LearnedResult", "PreparedGrids",
           "coarse_frequencies", "coarse_plan", "coarse_position", "features", "label_ranges",
           "prepare_grids",
           "window_geometry", "window_ranges"]
