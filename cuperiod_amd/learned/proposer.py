"""Learned candidate proposer plus fine refinement, with the plain search as fallback.

For each light curve on a linspace frequency grid (the plain, full-grid search):

1. a coarse BLS spectrum at L = 4,096 log-spaced frequencies spanning the grid, stacked
   over time chunks. A BLS peak of a curve spanning T days is about duration/T wide in
   relative frequency, far narrower than the coarse grid's spacing ``dlog`` for the
   baselines the plain grid is sized for, so a whole-curve coarse spectrum misses most
   peaks between its samples. The curve is cut into runs of points at most
   ``chunk_duration / dlog`` days long, each run is searched on the coarse grid (same
   engine, the plain search's segment durations, ``coarse_oversample`` bins per
   duration), and the chunks' SNR^2 are summed per coarse frequency: each chunk's peak
   spans about one coarse cell, and the stack keeps most of the transit's signal;
2. features: the stacked power standardised per curve and a 512-bin cadence histogram
   (``cpp_stack_features``);
3. the network scores every coarse cell as the one nearest the plain search's best
   frequency (it is trained on that cell +-1); greedy windows are taken around the most
   probable cells (``cpp_head_windows``), each claiming its centre +-``radius`` cells,
   and kept, most probable first, until their probability mass reaches ``window_mass``
   or ``max_windows`` are kept;
4. a curve whose kept mass is below ``threshold`` falls back to the full grid; every
   other curve is searched at the full grid's frequencies within ``fine_per_window / 2``
   grid steps of every cell its kept windows claim, and at least half a coarse cell
   each side of them, so a claimed cell that is the nearest one to the best frequency
   puts the best frequency inside the refined range on any grid.

Refinement evaluates the plain grid's own frequencies with the plain search's segment
binning (the full search's descriptors, cut to the windows), so a refined curve's best
period is the plain best period whenever a window covers it. Refined and fallback curves
share one engine launch.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Sequence

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.batch import engine
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.learned.network import L, ProposerNet, load_parameters
from cuperiod_amd.methods.bls import _log_segments

#: Weights trained on synthetic gapped curves labelled by the plain search
#: (``cuperiod_amd.learned.train``).
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "weights" / "proposer_gapped.npz"
#: Stacked coarse power is standardised per curve and clipped to +-FEATURE_CLIP.
FEATURE_CLIP = 30.0
#: Chunks with fewer points are left out of the coarse stack.
MIN_CHUNK_POINTS = 16


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


def label_cells(freq: np.ndarray, f_lo: np.ndarray, dlog: np.ndarray,
                radius: int = 1) -> np.ndarray:
    """``[n, 2]`` training targets: the coarse cell nearest each frequency, +-``radius``
    cells (clipped to the grid)."""
    near = np.rint(np.log(freq / f_lo) / dlog)
    out = np.stack([np.clip(near - radius, 0, L - 1), np.clip(near + radius, 0, L - 1)], axis=1)
    return out.astype(np.int32)


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


def time_chunks(cs: engine.CurveSet, max_span: np.ndarray):
    """``(parent, start, count)`` of the runs of each curve's points: consecutive points
    spanning at most ``max_span[c]`` days, runs of fewer than MIN_CHUNK_POINTS dropped.
    A curve whose times are not sorted is one run."""
    parents, starts, counts = [], [], []
    for c, t in enumerate(cs.times):
        tau = t - t[0]
        if tau.size > 1 and np.any(np.diff(tau) < 0):
            parents.append(np.array([c], dtype=np.int64))
            starts.append(np.array([0], dtype=np.int64))
            counts.append(np.array([tau.size], dtype=np.int64))
            continue
        edges = [0]
        while edges[-1] < tau.size:
            edges.append(int(np.searchsorted(tau, tau[edges[-1]] + max_span[c], side="right")))
        e = np.asarray(edges, dtype=np.int64)
        n = np.diff(e)
        keep = n >= MIN_CHUNK_POINTS
        parents.append(np.full(int(keep.sum()), c, dtype=np.int64))
        starts.append(e[:-1][keep])
        counts.append(n[keep])
    return np.concatenate(parents), np.concatenate(starts), np.concatenate(counts)


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
    """What depends only on the curves and grids: the plain search's packed batch (the
    fallback, and the descriptors refinement is cut from), the chunked coarse search and
    the window geometry."""

    cs: engine.CurveSet
    grids: list
    full: engine.PackedBatch
    coarse: engine.PackedBatch
    chunk_start: object = None
    chunk_base: object = None
    geometry: tuple = ()
    geometry_host: tuple = ()
    n_chunks: int = 0


def prepare_grids(cs: engine.CurveSet, grids: Sequence[engine.LinspaceGrid],
                  settings: BLSSettings, *, full: engine.PackedBatch | None = None,
                  fine_per_window: int = 2000, chunk_duration: float = 0.2,
                  coarse_oversample: int = 3) -> PreparedGrids:
    """The plain search's packed batch (built here unless given), the chunked coarse
    search's packed descriptors and the window geometry of a batch."""
    import torch

    grids = list(grids)
    if full is None:
        full = engine.pack_plans(cs, engine.plan_grids(cs, grids, settings), settings)
    f_lo, dlog, half = window_geometry(grids, fine_per_window)
    parent, start, count = time_chunks(cs, chunk_duration / dlog)
    chunks = engine.sub_curves(cs, parent, start, count)
    coarse_settings = settings.model_copy(update={"bins_per_duration": int(coarse_oversample)})
    # one coarse plan per distinct grid, shared by its curves' chunks (the packer then
    # sends each plan's periods once)
    by_grid: dict[tuple[float, float], engine.CurvePlan] = {}
    plans = []
    for c, g in enumerate(grids):
        key = (float(g.start), float(g.stop))
        if key not in by_grid:
            by_grid[key] = coarse_plan(c, g, settings)
        plans.append(by_grid[key])
    coarse = engine.pack_plans(chunks, [plans[p] for p in parent.tolist()], coarse_settings)
    chunk_start = np.searchsorted(parent, np.arange(len(grids) + 1)).astype(np.int32)
    dev = cs.tau.device
    # windows claim their centre +-radius cells (no claim width); refinement adds the
    # half-width around the claimed cells (refine_ranges)
    geometry = tuple(torch.from_numpy(a).to(dev) for a in (f_lo, dlog, np.zeros_like(half)))
    return PreparedGrids(cs=cs, grids=grids, full=full, coarse=coarse,
                         chunk_start=torch.from_numpy(chunk_start).to(dev),
                         chunk_base=coarse.curve_base_d, geometry=geometry,
                         geometry_host=(f_lo, dlog, half), n_chunks=int(parent.size))


def stacked_features(prep: PreparedGrids, out):
    """``[n, L, 2]`` FP32 device features from a coarse search's outputs."""
    import torch

    cs = prep.cs
    device = out.device
    n = len(prep.grids)
    x = torch.empty((n, L, 2), dtype=torch.float32, device=device)
    pbase = torch.from_numpy(cs.point_base).to(device)
    npts = torch.from_numpy(cs.n_points.astype(np.int32)).to(device)
    base = torch.from_numpy(cs.baseline).to(device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    lib = _lib.load_library()
    st = lib.cpp_stack_features(stream, ctypes.c_void_p(out.data_ptr()),
                                ctypes.c_void_p(prep.chunk_start.data_ptr()),
                                ctypes.c_void_p(prep.chunk_base.data_ptr()),
                                ctypes.c_void_p(cs.tau.data_ptr()),
                                ctypes.c_void_p(pbase.data_ptr()),
                                ctypes.c_void_p(npts.data_ptr()),
                                ctypes.c_void_p(base.data_ptr()), n, FEATURE_CLIP,
                                ctypes.c_void_p(x.data_ptr()))
    if st != _lib.CPB_OK:
        raise _lib.status_error(st, f"proposer features: status {st}")
    return x


def coarse_features(prep: PreparedGrids):
    """Run the chunked coarse search and return the network's input features."""
    out, _ = engine.run_packed(prep.coarse)
    return stacked_features(prep, out)


def refine_ranges(prep: PreparedGrids, span_lo: np.ndarray, span_hi: np.ndarray,
                  curve: np.ndarray):
    """Ascending-period index ranges ``[k0, k1)`` of the full grid of ``curve`` that hold
    every frequency within the refinement half-width of the cells ``span_lo ..
    span_hi`` and every frequency whose nearest coarse cell is one of them (one window
    per row), so the probability a window claims is always searched."""
    f_lo, dlog, half = (a[curve] for a in prep.geometry_host)
    start = np.array([g.start for g in prep.grids])[curve]
    stop = np.array([g.stop for g in prep.grids])[curve]
    num = np.array([g.num for g in prep.grids], dtype=np.int64)[curve]
    step = (stop - start) / (num - 1)
    # every frequency within the half-width of a claimed cell, and every frequency whose
    # nearest coarse cell is claimed (half a cell each side), whichever reaches further
    fa = np.minimum(f_lo * np.exp(span_lo * dlog) - half, f_lo * np.exp((span_lo - 0.5) * dlog))
    fb = np.maximum(f_lo * np.exp(span_hi * dlog) + half, f_lo * np.exp((span_hi + 0.5) * dlog))
    i_lo = np.clip(np.floor((fa - start) / step) - 1, 0, num - 1).astype(np.int64)
    i_hi = np.clip(np.ceil((fb - start) / step) + 1, 0, num - 1).astype(np.int64)
    return num - 1 - i_hi, num - i_lo


class LearnedProposer:
    """Chunked coarse search, network windows, refinement and fallback over a batch.

    The default ``window_mass`` (0.99) and ``threshold`` (0.5) were chosen on 2,048
    held-out gapped curves (a seed apart from the training and benchmark curves): from
    mass 0.99 up every setting of the threshold found the plain best period on every
    curve, and 0.5 still sends the least confident curves (0.05% there) to the full grid.
    """

    def __init__(self, params: dict[str, np.ndarray] | None = None, *, max_windows: int = 32,
                 radius: int = 1, fine_per_window: int = 2000, window_mass: float = 0.99,
                 threshold: float = 0.5, chunk_duration: float = 0.2,
                 coarse_oversample: int = 3) -> None:
        if params is None:
            params = load_parameters(DEFAULT_WEIGHTS)
        self.net = ProposerNet(params)
        self.max_windows = int(max_windows)
        self.radius = int(radius)
        self.fine_per_window = int(fine_per_window)
        self.window_mass = float(window_mass)
        self.threshold = float(threshold)
        self.chunk_duration = float(chunk_duration)
        self.coarse_oversample = int(coarse_oversample)

    def prepare(self, cs: engine.CurveSet, grids: Sequence[engine.LinspaceGrid],
                settings: BLSSettings, full: engine.PackedBatch | None = None) -> PreparedGrids:
        """Plans that do not depend on the network's output (built once per batch)."""
        return prepare_grids(cs, grids, settings, full=full,
                             fine_per_window=self.fine_per_window,
                             chunk_duration=self.chunk_duration,
                             coarse_oversample=self.coarse_oversample)

    def propose(self, prep: PreparedGrids):
        """Coarse search, features and windows: ``(centre, mass, span)`` host arrays."""
        x = coarse_features(prep)
        centre, mass, span = self.net.windows(x, self.max_windows, self.radius, *prep.geometry)
        return centre.cpu().numpy(), mass.cpu().numpy(), span.cpu().numpy()

    def select(self, centre: np.ndarray, mass: np.ndarray):
        """Per window, whether it is kept (most probable first until ``window_mass``), and
        each curve's kept mass."""
        m = np.where(centre >= 0, mass, 0.0).astype(np.float64)
        order = np.argsort(-m, axis=1, kind="stable")
        ms = np.take_along_axis(m, order, axis=1)
        cum = np.cumsum(ms, axis=1)
        # windows before the one that reaches window_mass, plus that one
        before = np.concatenate([np.zeros((m.shape[0], 1)), cum[:, :-1]], axis=1)
        keep_sorted = (before < self.window_mass) & (ms > 0.0)
        keep = np.zeros_like(keep_sorted)
        np.put_along_axis(keep, order, keep_sorted, axis=1)
        conf = np.where(keep, m, 0.0).sum(axis=1)
        return keep, conf

    def search(self, prep: PreparedGrids, settings: BLSSettings) -> LearnedResult:
        """Learned proposal + refinement, fallback below ``threshold``."""
        import torch

        t0 = time.perf_counter()
        centre, mass, span = self.propose(prep)
        t1 = time.perf_counter()
        keep, conf = self.select(centre, mass)
        fallback = conf < self.threshold
        n = len(prep.grids)
        rows, cols = np.nonzero(keep & ~fallback[:, None])
        k0, k1 = refine_ranges(prep, span[rows, cols, 0], span[rows, cols, 1], rows)
        fb = np.flatnonzero(fallback)
        curve = np.concatenate([rows, fb])
        k0 = np.concatenate([k0, np.zeros(fb.size, dtype=np.int64)])
        k1 = np.concatenate([k1, prep.full.curve_count[fb]])
        curve, k0, k1 = _merge_ranges(curve, k0, k1)
        wb = engine.window_batch(prep.full, curve, k0, k1)
        t2 = time.perf_counter()
        out, summary = engine.run_packed(wb)
        s = summary.cpu().numpy()
        torch.cuda.synchronize()
        best_col = wb.curve_base + s[:, 2].astype(np.int64)
        r = np.searchsorted(wb.segment_table["out_offset"][:wb.n_segments], best_col,
                            side="right") - 1
        full_col = wb.source_col[r] + (best_col - wb.segment_table["out_offset"][r])
        best_period = engine.periods_of_columns(prep.full, full_col)
        best_period = np.where(wb.curve_count > 0, best_period, np.nan)
        t3 = time.perf_counter()
        return LearnedResult(
            best_period=best_period, best_power=s[:, 3].copy(), fallback=fallback,
            confidence=conf, n_fine=wb.curve_count.copy(),
            timings={"propose": t1 - t0, "plan": t2 - t1, "refine": t3 - t2},
        )


def _merge_ranges(curve: np.ndarray, k0: np.ndarray, k1: np.ndarray):
    """Per curve, the union of ranges ``[k0, k1)`` as disjoint sorted ranges:
    ``(curve, k0, k1)``."""
    order = np.lexsort((k0, curve))
    curve, k0, k1 = curve[order], k0[order], k1[order]
    if curve.size == 0:
        return curve, k0, k1
    # running maximum of the range ends within each curve
    ends = k1.copy()
    start_of = np.concatenate([[True], curve[1:] != curve[:-1]])
    grp = np.cumsum(start_of) - 1
    for_max = ends + grp.astype(np.int64) * (int(k1.max()) + 1)
    run_max = np.maximum.accumulate(for_max) - grp.astype(np.int64) * (int(k1.max()) + 1)
    prev_end = np.concatenate([[-1], run_max[:-1]])
    new = start_of | (k0 > prev_end)
    idx = np.flatnonzero(new)
    last = np.concatenate([idx[1:], [curve.size]]) - 1
    return curve[idx], k0[idx], run_max[last]


__all__ = ["DEFAULT_WEIGHTS", "LearnedProposer", "LearnedResult", "PreparedGrids",
           "coarse_features", "coarse_frequencies", "coarse_plan", "coarse_position",
           "label_cells", "label_ranges", "prepare_grids", "refine_ranges", "stacked_features",
           "time_chunks", "window_geometry"]
