"""Batched BLS over many light curves in one launch sequence.

cuPeriod searches a batch one light curve at a time, each with its own uploads, launches
and a (7, F) download (src/cuperiod/batch/runner.py:214-250,
src/cuperiod/methods/_bls_core.py:741-840). Here a chunk of curves is packed into one
set of device arrays and searched by one work list of (curve, segment, period tile)
items; each curve's spectrum stays on the device, where ``cpb_bls_summarize`` reduces
it to the SDE statistics and the best period. Only full spectra that a caller asks for
come back to the host.

Grids are described parametrically so that no per-curve period array is built or sent:

* :class:`LinspaceGrid` -- ``np.linspace(start, stop, num)`` frequencies (a custom
  uniform frequency grid, split into log period segments as ``_log_segments`` does,
  src/cuperiod/methods/bls.py:128-152);
* ``None`` -- the curve's default BLS grid (``_segment_periods``, ``:71-105``): one
  ``np.arange`` of frequencies per log segment;
* a :class:`~cuperiod_amd.core.grid.GridSpec` -- explicit periods (sent as values).

The device regenerates every parametric frequency with numpy's own recipe, so the
searched periods are bit for bit those of the materialised grid.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Any, Sequence

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.core.errors import BackendUnavailableError, InsufficientDataError
from cuperiod_amd.core.grid import GridSpec, check_grid_size, linspace_frequencies
from cuperiod_amd.methods import _bls_core as core
from cuperiod_amd.methods.bls import (
    _grid_segments,
    _log_segment_index,
    _segment_bounds,
    _segment_durations,
)


@dataclass(frozen=True)
class LinspaceGrid:
    """Frequencies ``np.linspace(start, stop, num)`` (cycles/day), ``0 < start < stop``."""

    start: float
    stop: float
    num: int

    def __post_init__(self) -> None:
        if not (math.isfinite(self.start) and math.isfinite(self.stop)):
            raise ValueError("LinspaceGrid bounds must be finite")
        if not 0.0 < self.start < self.stop:
            raise ValueError("LinspaceGrid needs 0 < start < stop (finite periods)")
        if int(self.num) < 2:
            raise ValueError("LinspaceGrid needs num >= 2")
        check_grid_size(int(self.num), context="LinspaceGrid")

    @property
    def size(self) -> int:
        return int(self.num)

    def frequencies(self) -> np.ndarray:
        return linspace_frequencies(self.start, self.stop, self.num)

    def periods(self) -> np.ndarray:
        """Ascending periods ``1/f[::-1]`` (what ``GridSpec.period`` gives)."""
        return np.ascontiguousarray(1.0 / self.frequencies()[::-1])

    def as_gridspec(self) -> GridSpec:
        return GridSpec(kind="frequency", values=self.frequencies())

    def period_at(self, k: np.ndarray | int) -> np.ndarray:
        """Period of ascending-period index ``k`` by the device's recipe."""
        k = np.asarray(k, dtype=np.int64)
        i = self.num - 1 - k
        step = (float(self.stop) - float(self.start)) / (self.num - 1)
        f = np.where(i == self.num - 1, float(self.stop), i.astype(np.float64) * step + float(self.start))
        return 1.0 / f


@dataclass
class _SegPlan:
    """One segment of one curve: binning plus where its periods come from."""

    curve: int
    n_periods: int
    durations: np.ndarray
    p_max: float
    mode: int
    lin_start: float = 0.0
    lin_step: float = 0.0
    lin_stop: float = 0.0
    lin_num: int = 0
    lin_k0: int = 0
    periods: np.ndarray | None = None
    # the reference segment of these periods (the engine takes the band of its peak
    # rounding per reference segment); -1: the segment's position in its plan
    group: int = -1
    # cuts of a parametric segment where its phase span crosses the engine's span limits,
    # by oversample: they depend only on the fields above, so every batch that packs this
    # segment (the plain search, a learned search's fallback) computes them once
    cuts: dict = field(default_factory=dict, repr=False, compare=False)


@dataclass
class CurvePlan:
    """The segments of one curve and a way to name its periods."""

    segments: list[_SegPlan] = field(default_factory=list)
    n_periods: int = 0

    def period_at(self, k: int) -> float:
        """Period of ascending index ``k`` of this curve's grid."""
        for sg in self.segments:
            if k < sg.n_periods:
                if sg.mode == 0:
                    return float(sg.periods[k])
                i = sg.lin_num - 1 - (sg.lin_k0 + k)
                if sg.mode == 1 and i == sg.lin_num - 1:
                    f = sg.lin_stop
                elif sg.mode == 2 and i == 0:
                    f = sg.lin_start
                else:
                    f = float(i) * sg.lin_step + sg.lin_start
                return 1.0 / f
            k -= sg.n_periods
        raise IndexError("period index out of range")

    def periods_at(self, k: np.ndarray) -> np.ndarray:
        """Periods of ascending indices ``k`` (an int array), by the device's recipe:
        each index's segment is found by bisection and the parametric periods are formed
        in one vectorised pass."""
        k = np.asarray(k, dtype=np.int64)
        sizes = np.array([sg.n_periods for sg in self.segments], dtype=np.int64)
        ends = np.cumsum(sizes)
        if np.any((k < 0) | (k >= (int(ends[-1]) if ends.size else 0))):
            raise IndexError("period index out of range")
        out = np.empty(k.shape, dtype=np.float64)
        if k.size == 0:
            return out
        sid = np.searchsorted(ends, k, side="right")
        segs = self.segments
        modes = np.array([sg.mode for sg in segs], dtype=np.int64)
        # grid index i of each period: num - 1 - (k0 + k - segment start), parametric ones
        top = np.array([sg.lin_num - 1 - sg.lin_k0 for sg in segs], dtype=np.int64)
        top += ends - sizes
        i = np.take(top, sid) - k
        start = np.take(np.array([sg.lin_start for sg in segs], dtype=np.float64), sid)
        f = i.astype(np.float64) * np.take(np.array([sg.lin_step for sg in segs]), sid) + start
        # the grid's end samples are its exact bounds (mode 1: i = num - 1 is stop; mode 2:
        # i = 0 is start)
        last = np.take(np.array([sg.lin_num - 1 for sg in segs], dtype=np.int64), sid)
        for j in np.flatnonzero((i == 0) | (i == last)).tolist():
            s = int(sid[j])
            if modes[s] == 1 and i[j] == last[j]:
                f[j] = segs[s].lin_stop
            elif modes[s] == 2 and i[j] == 0:
                f[j] = segs[s].lin_start
        with np.errstate(divide="ignore"):  # explicit segments' f is 0, replaced below
            np.divide(1.0, f, out=out)
        if (modes == 0).any():
            explicit = np.take(modes, sid) == 0
            for s in np.unique(sid[explicit]).tolist():
                sel = sid == s
                out[sel] = segs[s].periods[k[sel] - int(ends[s] - sizes[s])]
        return out

    def periods(self) -> np.ndarray:
        """Every period of the curve's grid, ascending (materialised on request)."""
        parts = []
        for sg in self.segments:
            if sg.mode == 0:
                parts.append(sg.periods)
                continue
            k = np.arange(sg.n_periods, dtype=np.int64)
            i = sg.lin_num - 1 - (sg.lin_k0 + k)
            f = i.astype(np.float64) * sg.lin_step + sg.lin_start
            if sg.mode == 1:
                f = np.where(i == sg.lin_num - 1, sg.lin_stop, f)
            else:
                f = np.where(i == 0, sg.lin_start, f)
            parts.append(1.0 / f)
        return np.concatenate(parts) if parts else np.zeros(0)


def _arange_params(start: float, stop: float, step: float) -> tuple[float, int]:
    """``(step', n_kept)`` of ``np.arange(start, stop, step)`` restricted to values
    ``< stop``: numpy fills ``start + i*step'`` with ``step' = (start+step) - start``."""
    n = max(0, math.ceil((stop - start) / step))
    if n == 0:
        return step, 0
    step2 = (start + step) - start
    while n > 0:
        last = start if n == 1 else start + float(n - 1) * step2
        if last < stop:
            break
        n -= 1
    return step2, n


def _default_plan(curve: int, baseline: float, time: np.ndarray, settings: BLSSettings) -> CurvePlan:
    """The default grid of a curve (``_segment_periods``) as arange segments."""
    bounds = _segment_bounds(baseline, settings)
    plan = CurvePlan()
    if not bounds:
        return plan
    df = settings.grid_duration_frac / (settings.grid_oversample * baseline)
    total = sum(max(0, math.ceil((1.0 / lo - 1.0 / hi) / df)) for lo, hi in bounds)
    check_grid_size(total, automatic=True, time=time, baseline=baseline,
                    context="BLS default grid")
    for p_lo, p_hi in bounds:
        start, stop = 1.0 / p_hi, 1.0 / p_lo
        step2, n = _arange_params(start, stop, df)
        if n == 0:
            continue
        plan.segments.append(
            _SegPlan(curve=curve, n_periods=n, durations=_segment_durations(p_lo, settings),
                     p_max=1.0 / start, mode=2, lin_start=start, lin_step=step2,
                     lin_num=n, lin_k0=0)
        )
        plan.n_periods += n
    return plan


def _linspace_plan(curve: int, grid: LinspaceGrid, settings: BLSSettings) -> CurvePlan:
    """A linspace frequency grid split into log period segments from its shortest
    period, as ``_log_segments`` splits the materialised grid."""
    return _linspace_plans([curve], [grid], settings)[0]


def _linspace_periods(start: np.ndarray, stop: np.ndarray, num: np.ndarray,
                      k: np.ndarray) -> np.ndarray:
    """Ascending-period index ``k`` of linspace grids, by numpy's (and the device's)
    recipe: ``f_i = i*step + start`` with ``f_{num-1} = stop``, ``i = num-1-k``."""
    i = num - 1 - k
    step = (stop - start) / (num - 1).astype(np.float64)
    f = i.astype(np.float64) * step + start
    return 1.0 / np.where(i == num - 1, stop, f)


def _linspace_plans(curves: Sequence[int], grids: Sequence[LinspaceGrid],
                    settings: BLSSettings) -> list[CurvePlan]:
    """:func:`_linspace_plan` of many grids at once: every segment boundary is found by
    a bisection on the period index run for all (grid, segment) pairs together; the
    segment index of a period is ``_log_segment_index``'s array expression."""
    n = len(grids)
    if n == 0:
        return []
    start = np.array([g.start for g in grids], dtype=np.float64)
    stop = np.array([g.stop for g in grids], dtype=np.float64)
    num = np.array([g.num for g in grids], dtype=np.int64)
    p0 = 1.0 / stop  # period of ascending index 0 (f = stop)
    log_factor = math.log(settings.segment_factor)

    def seg_index(rows: np.ndarray, k: np.ndarray) -> np.ndarray:
        per = _linspace_periods(start[rows], stop[rows], num[rows], k)
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.floor(np.log(per / p0[rows]) / log_factor + 1e-9).astype(np.int64)

    rows = np.arange(n)
    last = seg_index(rows, num - 1)
    # (grid, s) pairs for s = 1 .. last: the first ascending index with index >= s
    pair_rows = np.repeat(rows, last)
    pair_s = np.concatenate([np.arange(1, m + 1) for m in last]) if last.sum() else \
        np.zeros(0, dtype=np.int64)
    lo = np.zeros(pair_rows.size, dtype=np.int64)
    hi = num[pair_rows] - 1
    while np.any(lo < hi):
        mid = (lo + hi) // 2
        ge = seg_index(pair_rows, mid) >= pair_s
        open_ = lo < hi
        hi = np.where(open_ & ge, mid, hi)
        lo = np.where(open_ & ~ge, mid + 1, lo)
    firsts = np.split(lo, np.cumsum(last)[:-1]) if n > 1 else [lo]
    # every segment's longest period (its last index), all grids at once
    seg_rows = np.repeat(rows, last + 1)
    seg_end = np.concatenate([np.concatenate([f, [num[r]]]) for r, f in enumerate(firsts)])
    p_last = _linspace_periods(start[seg_rows], stop[seg_rows], num[seg_rows],
                               np.maximum(seg_end - 1, 0))
    p_last_of = np.split(p_last, np.cumsum(last + 1)[:-1]) if n > 1 else [p_last]
    durations: dict[float, np.ndarray] = {}
    plans: list[CurvePlan] = []
    for r in range(n):
        grid = grids[r]
        g_num = int(num[r])
        step = (float(grid.stop) - float(grid.start)) / (g_num - 1)
        # segment s covers [first(s), first(s+1)); indices never reached are skipped, as
        # np.unique skips them in _log_segments
        bounds = [0, *firsts[r].tolist(), g_num]
        plan = CurvePlan()
        p_ends = p_last_of[r].tolist()
        for s, k0, k1 in zip(range(int(last[r]) + 1), bounds[:-1], bounds[1:]):
            if k1 <= k0:
                continue
            p_lo = float(p0[r]) * settings.segment_factor**s
            dur = durations.get(p_lo)
            if dur is None:
                dur = durations[p_lo] = _segment_durations(p_lo, settings)
            p_max = p_ends[s]
            plan.segments.append(
                _SegPlan(curve=curves[r], n_periods=k1 - k0, durations=dur, p_max=p_max,
                         mode=1, lin_start=float(grid.start), lin_step=step,
                         lin_stop=float(grid.stop), lin_num=g_num, lin_k0=k0))
            plan.n_periods += k1 - k0
        plans.append(plan)
    return plans


def _explicit_plan(curve: int, grid: GridSpec, settings: BLSSettings) -> CurvePlan:
    plan = CurvePlan()
    for periods, durations in _grid_segments(grid, settings):
        if periods.size == 0:
            continue
        plan.segments.append(
            _SegPlan(curve=curve, n_periods=int(periods.size), durations=durations,
                     p_max=float(periods.max()), mode=0,
                     periods=np.ascontiguousarray(periods, dtype=np.float64))
        )
        plan.n_periods += int(periods.size)
    return plan


@dataclass
class BatchResult:
    """Per-curve outcome of :func:`bls_batch` (arrays aligned with the input curves).

    ``best_period``/``best_power`` follow ``Periodogram.best_period``; ``mean``/``std``
    are the SDE statistics of the sanitised power; ``spectra[c]`` (when requested) holds
    the curve's seven output rows in ascending-period order.
    """

    best_period: np.ndarray
    best_power: np.ndarray
    best_index: np.ndarray
    mean: np.ndarray
    std: np.ndarray
    n_periods: np.ndarray
    plans: list[CurvePlan]
    spectra: list[dict[str, np.ndarray]] | None = None


@dataclass
class CurveSet:
    """A chunk's light curves on the device: FP64 ``tau = t - t_min``, ``y*ivar`` and
    ``ivar`` concatenated, plus the per-curve descriptors the engine reads."""

    tau: Any
    yw: Any
    ivar: Any
    total_points: int
    curves: Any
    curves32: Any
    t_min: np.ndarray
    baseline: np.ndarray
    n_points: np.ndarray
    point_base: np.ndarray
    times: list[np.ndarray]
    # the CURVE_DTYPE tables behind ``curves`` / ``curves32`` (kept alive with them)
    tables: tuple = ()
    # every point's ivar on the host, concatenated (sums over runs of points)
    host_ivar: np.ndarray | None = None


@dataclass
class PackedBatch:
    """Device-resident inputs and descriptors of one chunk of curves."""

    tau: Any
    yw: Any
    ivar: Any
    total_points: int
    curves: Any
    segments: Any
    n_segments: int
    periods: Any
    dur_bins: Any
    n_out: int
    curve_base: np.ndarray
    curve_count: np.ndarray
    curve_base_d: Any
    curve_count_d: Any
    plans: list[CurvePlan]
    oversample: int
    obj_flag: int
    curve_set: CurveSet | None = None
    accumulation: int = 0
    # the SEG_DTYPE table behind ``segments``
    segment_table: np.ndarray | None = None
    # explicit periods on the host (what ``periods`` holds on the device)
    periods_host: np.ndarray | None = None
    # window batches: the full batch's column of each piece's first period
    source_col: np.ndarray | None = None


def _device():
    import torch

    ctx = _lib.context(0)
    if not torch.cuda.is_available():
        raise BackendUnavailableError("BLS batch: torch sees no HIP device")
    return torch.device("cuda", ctx.device)


_UPLOAD_STREAMS: dict[int, Any] = {}


def _upload(a: np.ndarray, dtype, device) -> Any:
    """``a`` on the device. The copy runs on an upload stream of its own (default
    priority), so packing the next batch does not wait for the search still running on
    torch's current stream; that stream waits for the copy before any later work, and
    the tensor is marked as used there for the allocator."""
    import torch

    a = np.ascontiguousarray(a, dtype=dtype)
    if a.size == 0:
        a = np.zeros(1, dtype=dtype)
    device = torch.device(device)
    up = _UPLOAD_STREAMS.get(device.index)
    if up is None:
        up = _UPLOAD_STREAMS[device.index] = torch.cuda.Stream(device)
    with torch.cuda.stream(up):
        t = torch.from_numpy(a).to(device)
        copied = torch.cuda.Event()
        copied.record(up)
    current = torch.cuda.current_stream(device)
    current.wait_event(copied)
    t.record_stream(current)
    return t


#: numpy layouts of the descriptors (field for field the ctypes structures, which
#: mirror cpb_curve / cpb_segment): descriptor tables are filled column-wise and handed
#: to the library as ctypes arrays over the same memory.
CURVE_DTYPE = np.dtype(_lib.Curve)
SEG_DTYPE = np.dtype(_lib.Segment)


def curve_array(table: np.ndarray):
    """A ctypes ``Curve`` array over a ``CURVE_DTYPE`` table (the table must stay alive)."""
    return (_lib.Curve * table.size).from_buffer(table)


def pack_curves(lcs: Sequence[tuple[np.ndarray, np.ndarray, np.ndarray]]) -> CurveSet:
    """Host preparation (median-centred flux, ``ivar``, time origin, floor) and one
    upload of a chunk's curves ``(t, flux, dy)``."""
    device = _device()
    taus, yws, ivs, times = [], [], [], []
    n = len(lcs)
    table = np.zeros(max(1, n), dtype=CURVE_DTYPE)
    table32 = np.zeros(max(1, n), dtype=CURVE_DTYPE)
    t_min = np.zeros(n)
    baseline = np.zeros(n)
    n_points = np.zeros(n, dtype=np.int64)
    floor64 = np.zeros(n)
    floor32 = np.zeros(n)
    sum_y = np.zeros(n)
    sum_iv = np.zeros(n)
    for c, (t, y, dy) in enumerate(lcs):
        t_h, y_h, iv_h = core._host_inputs(t, y, dy)
        if t_h.size == 0:
            raise InsufficientDataError("BLS batch: empty light curve")
        t_min[c] = float(t_h.min())
        baseline[c] = float(t_h.max() - t_min[c])
        taus.append(t_h - t_min[c])
        yws.append(y_h * iv_h)
        ivs.append(iv_h)
        times.append(t_h)
        n_points[c] = t_h.size
        sum_iv[c] = float(iv_h.sum())
        # the reference's host sums (numpy), read by the reference rounding of the peak
        sum_y[c] = float(yws[-1].sum())
        floor64[c] = core._ivar_floor(int(t_h.size), sum_iv[c])
        floor32[c] = core._ivar_floor(int(t_h.size), sum_iv[c], float32=True)
    point_base = np.concatenate([[0], np.cumsum(n_points)[:-1]]).astype(np.int64) if n else \
        np.zeros(0, dtype=np.int64)
    for tab, floor in ((table, floor64), (table32, floor32)):
        tab["point_offset"][:n] = point_base
        tab["n_points"][:n] = n_points
        tab["t_min"][:n] = t_min
        tab["ivar_floor"][:n] = floor
        tab["sum_y"][:n] = sum_y
        tab["sum_ivar"][:n] = sum_iv
    off = int(n_points.sum())
    return CurveSet(
        tau=_upload(np.concatenate(taus) if taus else np.zeros(0), np.float64, device),
        yw=_upload(np.concatenate(yws) if yws else np.zeros(0), np.float64, device),
        ivar=_upload(np.concatenate(ivs) if ivs else np.zeros(0), np.float64, device),
        total_points=off, curves=curve_array(table), curves32=curve_array(table32),
        t_min=t_min, baseline=baseline, n_points=n_points, point_base=point_base,
        times=times, tables=(table, table32),
        host_ivar=np.concatenate(ivs) if ivs else np.zeros(0),
    )


def sub_curves(cs: CurveSet, parent: np.ndarray, start: np.ndarray,
               count: np.ndarray) -> CurveSet:
    """Curves made of runs of ``cs``'s points, sharing its device arrays: curve ``j`` is
    points ``start[j] .. start[j] + count[j] - 1`` of curve ``parent[j]`` (runs must not
    overlap). Each is searched as those points alone (times measured from the parent's
    origin, the parent's centred flux: BLS powers do not depend on a flux offset); the
    reference rounding of the peak is not applied to them."""
    parent = np.asarray(parent, dtype=np.int64)
    start = np.asarray(start, dtype=np.int64)
    count = np.asarray(count, dtype=np.int64)
    n = parent.size
    off = cs.point_base[parent] + start
    csum = np.concatenate([[0.0], np.cumsum(cs.host_ivar)])
    sum_iv = csum[off + count] - csum[off]
    floor = 4.0 * count * core._F64_EPS * sum_iv
    floor = np.where(floor > 0.0, floor, np.finfo(np.float64).tiny)
    table = np.zeros(max(1, n), dtype=CURVE_DTYPE)
    table["point_offset"][:n] = off
    table["n_points"][:n] = count
    table["t_min"][:n] = cs.t_min[parent]
    table["ivar_floor"][:n] = floor
    # sum_y = sum_ivar = 0: no reference rounding of these curves' peaks
    curves = curve_array(table)
    return CurveSet(tau=cs.tau, yw=cs.yw, ivar=cs.ivar, total_points=cs.total_points,
                    curves=curves, curves32=curves, t_min=cs.t_min[parent],
                    baseline=cs.baseline[parent], n_points=count, point_base=off,
                    times=[], tables=(table,), host_ivar=None)


def plan_grid(cs: CurveSet, c: int, grid: LinspaceGrid | GridSpec | None,
              settings: BLSSettings) -> CurvePlan:
    """The segment plan of curve ``c`` on ``grid`` (``None``: its default grid)."""
    if grid is None:
        return _default_plan(c, float(cs.baseline[c]), cs.times[c], settings)
    if isinstance(grid, LinspaceGrid):
        return _linspace_plan(c, grid, settings)
    return _explicit_plan(c, grid, settings)


def plan_grids(cs: CurveSet, grids: Sequence[LinspaceGrid | GridSpec | None],
               settings: BLSSettings) -> list[CurvePlan]:
    """:func:`plan_grid` of every curve (linspace grids segmented together)."""
    plans: list[CurvePlan | None] = [None] * len(grids)
    lin = [c for c, g in enumerate(grids) if isinstance(g, LinspaceGrid)]
    for c, plan in zip(lin, _linspace_plans(lin, [grids[c] for c in lin], settings)):
        plans[c] = plan
    for c, g in enumerate(grids):
        if plans[c] is None:
            plans[c] = plan_grid(cs, c, g, settings)
    return plans


def _parametric_need(par: dict[str, np.ndarray], idx: np.ndarray, k: np.ndarray,
                     oversample: int) -> np.ndarray:
    """Phase bins ``ceil(P/bd) + os + 1`` of period ``k`` of parametric segment ``idx``
    (arrays; ``par`` holds the segments' parameters), the periods by the device's recipe
    (``periods_at``)."""
    num = par["num"][idx]
    i = num - 1 - (par["k0"][idx] + k)
    start = par["start"][idx]
    f = i.astype(np.float64) * par["step"][idx] + start
    mode = par["mode"][idx]
    f = np.where((mode == 1) & (i == num - 1), par["stop"][idx], f)
    f = np.where((mode == 2) & (i == 0), start, f)
    return np.ceil((1.0 / f) / par["bd"][idx]).astype(np.int64) + int(oversample) + 1


def _parametric_width_ranges(sgs: Sequence[_SegPlan], oversample: int) -> list[np.ndarray]:
    """``core._width_ranges`` of many parametric segments at once, per segment an int64
    array of ``(start, stop, width)`` rows: their periods ascend with the index, so the
    first index past each span limit is found by a vectorised bisection instead of
    materialising every period."""
    par = {
        "start": np.array([sg.lin_start for sg in sgs], dtype=np.float64),
        "step": np.array([sg.lin_step for sg in sgs], dtype=np.float64),
        "stop": np.array([sg.lin_stop for sg in sgs], dtype=np.float64),
        "num": np.array([sg.lin_num for sg in sgs], dtype=np.int64),
        "k0": np.array([sg.lin_k0 for sg in sgs], dtype=np.int64),
        "mode": np.array([sg.mode for sg in sgs], dtype=np.int64),
        "bd": np.array([float(sg.durations.min()) / oversample for sg in sgs]),
    }
    n_seg = len(sgs)
    limits = np.asarray(core._SPAN_LIMITS, dtype=np.int64)
    n = np.array([sg.n_periods for sg in sgs], dtype=np.int64)
    idx = np.repeat(np.arange(n_seg), limits.size)
    lim = np.tile(limits, n_seg)
    lo = np.zeros(idx.size, dtype=np.int64)
    hi = n[idx].copy()
    while np.any(lo < hi):
        mid = (lo + hi) // 2
        past = _parametric_need(par, idx, np.minimum(mid, n[idx] - 1), oversample) > lim
        open_ = lo < hi
        hi = np.where(open_ & past, mid, hi)
        lo = np.where(open_ & ~past, mid + 1, lo)
    first = np.minimum(lo.reshape(n_seg, limits.size), n[:, None])
    # intervals [0, first_0), [first_0, first_1), ..., [first_last, n); empty ones dropped
    starts = np.concatenate([np.zeros((n_seg, 1), dtype=np.int64), first], axis=1).ravel()
    ends = np.concatenate([first, n[:, None]], axis=1).ravel()
    seg_of = np.repeat(np.arange(n_seg), limits.size + 1)
    keep = ends > starts
    starts, ends, seg_of = starts[keep], ends[keep], seg_of[keep]
    widths = _parametric_need(par, seg_of, ends - 1, oversample)
    counts = np.bincount(seg_of, minlength=n_seg)
    rows = np.stack([starts, ends, widths], axis=1).astype(np.int64)
    bounds = np.concatenate([[0], np.cumsum(counts)]).tolist()
    return [rows[a:b] for a, b in zip(bounds[:-1], bounds[1:])]


def pack_plans(cs: CurveSet, plans: Sequence[CurvePlan], settings: BLSSettings) -> PackedBatch:
    """Segment descriptors of ``plans`` (one per curve of ``cs``, in order) over the
    uploaded curves."""
    device = _device()
    oversample = int(settings.bins_per_duration)
    obj_flag = core._objective_flag(settings.objective)
    accumulation = core._accumulation_code(settings.accumulation, settings.precision)
    plans = list(plans)
    # descriptor pieces: a segment whose periods need different phase-bin layouts is cut
    # where the width crosses the engine's span limits (outputs do not depend on it);
    # per segment an int64 array of (start, stop, width) rows
    seg_rows: list[tuple[int, _SegPlan, np.ndarray | None]] = []
    group_of: dict[int, int] = {}
    bd_of: dict[int, float] = {}
    explicit_cuts: dict[int, np.ndarray] = {}
    todo: list[_SegPlan] = []
    for c, plan in enumerate(plans):
        for pos, sg in enumerate(plan.segments):
            group_of[id(sg)] = sg.group if sg.group >= 0 else pos
            bd = bd_of[id(sg)] = float(sg.durations.min()) / oversample
            width = int(np.ceil(sg.p_max / bd)) + oversample + 1
            rows = None
            if width <= core._SPAN_LIMITS[0] or sg.n_periods < 2:
                rows = np.array([[0, sg.n_periods, width]], dtype=np.int64)
            elif sg.mode == 0:
                rows = explicit_cuts.get(id(sg))
                if rows is None:
                    rows = explicit_cuts[id(sg)] = np.array(
                        core._width_ranges(sg.periods, bd, oversample), dtype=np.int64
                    ).reshape(-1, 3)
            elif oversample not in sg.cuts:
                todo.append(sg)
            seg_rows.append((c, sg, rows))
    if todo:
        for sg, rng in zip(todo, _parametric_width_ranges(todo, oversample)):
            sg.cuts[oversample] = rng
    seg_rows = [(c, sg, sg.cuts[oversample] if rows is None else rows)
                for c, sg, rows in seg_rows]
    dur_parts: list[np.ndarray] = []
    per_parts: list[np.ndarray] = []
    p_off = d_off = 0
    curve_count = np.array([plan.n_periods for plan in plans], dtype=np.int64)
    # per segment, once: its bin duration, box widths (deduplicated by value) and group
    seg_info: dict[int, tuple[float, int, int, int]] = {}
    dur_cache: dict[tuple[float, bytes], tuple[int, int]] = {}
    seg_list: list[_SegPlan] = []
    seg_pos: dict[int, int] = {}
    entry_pos = np.empty(len(seg_rows), dtype=np.int64)
    for e, (c, sg, rows) in enumerate(seg_rows):
        key = id(sg)
        pos = seg_pos.get(key)
        if pos is None:
            pos = seg_pos[key] = len(seg_list)
            seg_list.append(sg)
            bd = bd_of[key]
            dkey = (bd, sg.durations.tobytes())
            hit = dur_cache.get(dkey)
            if hit is None:
                db = np.asarray(core._duration_bins(sg.durations, bd), dtype=np.int32)
                hit = dur_cache[dkey] = (d_off, int(db.size))
                dur_parts.append(db)
                d_off += int(db.size)
            seg_info[key] = (bd, hit[0], hit[1], group_of[key])
        entry_pos[e] = pos
    counts = np.array([rows.shape[0] for _, _, rows in seg_rows], dtype=np.int64)
    n_pieces = int(counts.sum())
    table = np.zeros(max(1, n_pieces), dtype=SEG_DTYPE)
    if n_pieces:
        cuts = np.concatenate([rows for _, _, rows in seg_rows])
        col_a, col_b, col_w = cuts[:, 0], cuts[:, 1], cuts[:, 2]
        col_c = np.repeat(np.array([c for c, _, _ in seg_rows], dtype=np.int64), counts)
        col_s = np.repeat(entry_pos, counts)
    col_p = np.zeros(n_pieces, dtype=np.int64)
    # the same explicit periods (one array shared by several plans) go up once
    per_cache: dict[tuple[int, int, int], int] = {}
    first = np.concatenate([[0], np.cumsum(counts)[:-1]]) if counts.size else counts
    for e, (c, sg, rows) in enumerate(seg_rows):
        if sg.mode != 0:
            continue
        for r, (a, b, _) in enumerate(rows.tolist()):
            pkey = (id(sg.periods), a, b)
            off = per_cache.get(pkey)
            if off is None:
                off = per_cache[pkey] = p_off
                per_parts.append(sg.periods[a:b])
                p_off += b - a
            col_p[first[e] + r] = off
    if n_pieces:
        info = np.array([seg_info[id(sg)] for sg in seg_list], dtype=np.float64)
        lin = np.array([(sg.mode, sg.lin_start, sg.lin_step, sg.lin_stop, sg.lin_num,
                         sg.lin_k0) for sg in seg_list], dtype=np.float64)
        n_per = col_b - col_a
        out_off = np.concatenate([[0], np.cumsum(n_per)[:-1]])
        table["bin_duration"][:n_pieces] = info[col_s, 0]
        table["period_offset"][:n_pieces] = col_p
        table["n_periods"][:n_pieces] = n_per
        table["out_offset"][:n_pieces] = out_off
        table["width"][:n_pieces] = col_w
        table["dur_offset"][:n_pieces] = info[col_s, 1].astype(np.int64)
        table["n_dur"][:n_pieces] = info[col_s, 2].astype(np.int64)
        table["curve"][:n_pieces] = col_c
        table["group"][:n_pieces] = info[col_s, 3].astype(np.int64)
        table["mode"][:n_pieces] = lin[col_s, 0].astype(np.int64)
        table["lin_start"][:n_pieces] = lin[col_s, 1]
        table["lin_step"][:n_pieces] = lin[col_s, 2]
        table["lin_stop"][:n_pieces] = lin[col_s, 3]
        # lin_num and lin_k0 are integers below 2^53: exact through float64
        table["lin_num"][:n_pieces] = lin[col_s, 4].astype(np.int64)
        table["lin_k0"][:n_pieces] = lin[col_s, 5].astype(np.int64) + col_a
        out_off = int(n_per.sum())
    else:
        out_off = 0
    # a curve's first column: where its first piece starts (its outputs are contiguous)
    curve_base = np.concatenate([[0], np.cumsum(curve_count)[:-1]]).astype(np.int64) \
        if len(plans) else np.zeros(0, dtype=np.int64)
    segs = (_lib.Segment * table.size).from_buffer(table)
    return PackedBatch(
        tau=cs.tau, yw=cs.yw, ivar=cs.ivar, total_points=cs.total_points,
        curves=cs.curves32 if accumulation else cs.curves,
        segments=segs, n_segments=n_pieces, segment_table=table,
        periods_host=np.concatenate(per_parts) if per_parts else np.zeros(1),
        periods=_upload(np.concatenate(per_parts) if per_parts else np.zeros(1), np.float64, device),
        dur_bins=_upload(np.concatenate(dur_parts) if dur_parts else np.zeros(1), np.int32, device),
        n_out=out_off, curve_base=curve_base, curve_count=curve_count,
        curve_base_d=_upload(curve_base, np.int64, device),
        curve_count_d=_upload(curve_count, np.int64, device),
        plans=plans, oversample=oversample, obj_flag=obj_flag, curve_set=cs,
        accumulation=accumulation,
    )


def pack(
    lcs: Sequence[tuple[np.ndarray, np.ndarray, np.ndarray]],
    grids: Sequence[LinspaceGrid | GridSpec | None],
    settings: BLSSettings,
) -> PackedBatch:
    """Host preparation of a chunk: curves uploaded once, every curve's segment plan."""
    cs = pack_curves(lcs)
    return pack_plans(cs, plan_grids(cs, grids, settings), settings)


def window_batch(pb: PackedBatch, curve: np.ndarray, k0: np.ndarray,
                 k1: np.ndarray) -> PackedBatch:
    """The part of packed batch ``pb`` at ascending-period indices ``[k0, k1)`` of
    curve ``curve`` (one row per window; a curve's windows must not overlap), cut from
    its descriptor table: every kept period keeps its piece's binning, box widths and
    reference segment, so its outputs are those of the full search (up to the reference
    rounding of a segment's peak, which follows the periods searched). ``source_col``
    maps each new piece back to the full batch's columns."""
    import dataclasses as dc

    table = pb.segment_table[:pb.n_segments]
    p_start = table["out_offset"].astype(np.int64)
    p_end = p_start + table["n_periods"].astype(np.int64)
    g0 = pb.curve_base[curve] + np.asarray(k0, dtype=np.int64)
    g1 = pb.curve_base[curve] + np.asarray(k1, dtype=np.int64)
    ok = g1 > g0
    g0, g1 = g0[ok], g1[ok]
    order = np.argsort(g0, kind="stable")
    g0, g1 = g0[order], g1[order]
    first = np.searchsorted(p_start, g0, side="right") - 1
    last = np.searchsorted(p_start, g1 - 1, side="right") - 1
    cnt = last - first + 1
    w = np.repeat(np.arange(g0.size), cnt)
    piece = first[w] + (np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    lo = np.maximum(g0[w], p_start[piece])
    hi = np.minimum(g1[w], p_end[piece])
    keep = hi > lo
    piece, lo, hi = piece[keep], lo[keep], hi[keep]
    new = np.zeros(max(1, piece.size), dtype=SEG_DTYPE)
    m = piece.size
    if m:
        new[:m] = table[piece]
        shift = lo - p_start[piece]
        n_per = hi - lo
        explicit = new["mode"][:m] == 0
        new["period_offset"][:m] += np.where(explicit, shift, 0)
        new["lin_k0"][:m] += np.where(explicit, 0, shift)
        new["n_periods"][:m] = n_per
        new["out_offset"][:m] = np.concatenate([[0], np.cumsum(n_per)[:-1]])
    n_curves = len(pb.plans)
    counts = np.bincount(new["curve"][:m], weights=new["n_periods"][:m],
                         minlength=n_curves).astype(np.int64)
    bases = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    device = pb.curve_base_d.device
    return dc.replace(
        pb, segments=(_lib.Segment * new.size).from_buffer(new), n_segments=m,
        segment_table=new, n_out=int(counts.sum()), curve_base=bases, curve_count=counts,
        curve_base_d=_upload(bases, np.int64, device),
        curve_count_d=_upload(counts, np.int64, device), source_col=lo.astype(np.int64))


def periods_of_columns(pb: PackedBatch, cols: np.ndarray) -> np.ndarray:
    """Periods of output columns ``cols`` of a packed batch, by the device's recipes."""
    cols = np.asarray(cols, dtype=np.int64)
    table = pb.segment_table[:pb.n_segments]
    r = np.searchsorted(table["out_offset"], cols, side="right") - 1
    row = table[r]
    i_in = cols - row["out_offset"]
    out = np.empty(cols.shape, dtype=np.float64)
    explicit = row["mode"] == 0
    if explicit.any():
        out[explicit] = pb.periods_host[row["period_offset"][explicit] + i_in[explicit]]
    par = ~explicit
    if par.any():
        rw = row[par]
        num = rw["lin_num"].astype(np.int64)
        i = num - 1 - (rw["lin_k0"].astype(np.int64) + i_in[par])
        f = i.astype(np.float64) * rw["lin_step"] + rw["lin_start"]
        f = np.where((rw["mode"] == 1) & (i == num - 1), rw["lin_stop"], f)
        f = np.where((rw["mode"] == 2) & (i == 0), rw["lin_start"], f)
        out[par] = 1.0 / f
    return out


def window_plan(base: CurvePlan, ranges: Sequence[tuple[int, int]]) -> CurvePlan:
    """The part of ``base`` (a plan of parametric segments) at ascending-period indices
    in the disjoint, sorted ``ranges`` ``[k0, k1)``. Every kept period keeps its
    segment's durations and binning, so its outputs equal the full search's."""
    plan = CurvePlan()
    start = 0
    for pos, sg in enumerate(base.segments):
        if sg.mode == 0:
            raise ValueError("window_plan needs parametric segments")
        lo, hi = start, start + sg.n_periods
        for a, b in ranges:
            a, b = max(a, lo), min(b, hi)
            if a >= b:
                continue
            plan.segments.append(
                _SegPlan(curve=sg.curve, n_periods=b - a, durations=sg.durations, p_max=sg.p_max,
                         mode=sg.mode, lin_start=sg.lin_start, lin_step=sg.lin_step,
                         lin_stop=sg.lin_stop, lin_num=sg.lin_num, lin_k0=sg.lin_k0 + (a - lo),
                         group=sg.group if sg.group >= 0 else pos)
            )
            plan.n_periods += b - a
        start = hi
    return plan


def _vp(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def run_packed(pb: PackedBatch, *, out=None, summary=None):
    """Enqueue the batched search and the per-curve summaries on torch's current
    stream; returns ``(out, summary)`` device tensors (not synchronised)."""
    import torch

    ctx = _lib.context(0)
    device = torch.device("cuda", ctx.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    if out is None:
        out = torch.empty((_lib.N_OUTPUTS, max(1, pb.n_out)), dtype=torch.float64, device=device)
    if summary is None:
        summary = torch.empty((len(pb.plans), 4), dtype=torch.float64, device=device)
    with ctx.lock:
        ctx.check(ctx.lib.cpb_set_accumulation(ctx.handle, pb.accumulation), "accumulation")
        st = ctx.lib.cpb_bls_search_batch(
            ctx.handle, stream, _vp(pb.tau), _vp(pb.yw), _vp(pb.ivar), pb.total_points,
            pb.curves, len(pb.plans), _vp(pb.periods), _vp(pb.dur_bins), pb.segments,
            pb.n_segments, pb.oversample, pb.obj_flag, _vp(out), pb.n_out,
        )
        ctx.check(st, "batch search")
        st = ctx.lib.cpb_bls_summarize(
            ctx.handle, stream, _vp(out), _vp(pb.curve_base_d), _vp(pb.curve_count_d),
            len(pb.plans), _vp(summary),
        )
        ctx.check(st, "summarize")
    return out, summary


@dataclass
class PeakCandidates:
    """Device peak candidates of each curve (``cpb_peak_candidates``), on the host:
    ``index[c]`` ascending-frequency positions (sorted), ``values[c]`` their (power,
    depth, depth_snr, duration, transit_time), ``total[c]`` all candidates of the curve,
    ``written[c]`` those at or above the curve's threshold (``> cap``: truncated)."""

    index: list[np.ndarray]
    values: list[np.ndarray]
    total: np.ndarray
    written: np.ndarray
    cap: int


def peak_candidates(pb: PackedBatch, out, *, want: int = 512, cap: int = 1024) -> PeakCandidates:
    """Peak candidates of every curve of a finished (or enqueued) search, on the host."""
    import torch

    ctx = _lib.context(0)
    device = out.device
    n = len(pb.plans)
    cand = torch.empty((max(1, n), cap), dtype=torch.int32, device=device)
    counts = torch.empty((max(1, n), 2), dtype=torch.int64, device=device)
    values = torch.empty((max(1, n), cap, 5), dtype=torch.float64, device=device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    with ctx.lock:
        st = ctx.lib.cpb_peak_candidates(
            ctx.handle, stream, _vp(out), pb.n_out, _vp(pb.curve_base_d), _vp(pb.curve_count_d),
            n, int(want), int(cap), _vp(cand), _vp(counts), _vp(values),
        )
        ctx.check(st, "peak candidates")
    cand_h = cand.cpu().numpy()
    counts_h = counts.cpu().numpy()
    values_h = values.cpu().numpy()
    index, vals = [], []
    for c in range(n):
        m = int(min(counts_h[c, 1], cap))
        order = np.argsort(cand_h[c, :m], kind="stable")
        index.append(cand_h[c, :m][order].astype(np.int64))
        vals.append(values_h[c, :m][order])
    return PeakCandidates(index=index, values=vals, total=counts_h[:n, 0].copy(),
                          written=counts_h[:n, 1].copy(), cap=cap)


class BatchGraph:
    """A batched search and its per-curve summaries captured once into a HIP graph
    (``cpb_graph_create``); :meth:`launch` replays it on torch's current stream. The
    graph reads the packed inputs and writes ``out``/``summary`` in place, so ``pb``,
    ``out`` and ``summary`` are kept alive with it."""

    def __init__(self, pb: PackedBatch, out, summary) -> None:
        self.ctx = _lib.context(0)
        self.pb, self.out, self.summary = pb, out, summary
        handle = ctypes.c_void_p()
        with self.ctx.lock:
            self.ctx.check(self.ctx.lib.cpb_set_accumulation(self.ctx.handle, pb.accumulation),
                           "accumulation")
            st = self.ctx.lib.cpb_graph_create(
                self.ctx.handle, _vp(pb.tau), _vp(pb.yw), _vp(pb.ivar), pb.total_points,
                pb.curves, len(pb.plans), _vp(pb.periods), _vp(pb.dur_bins), pb.segments,
                pb.n_segments, pb.oversample, pb.obj_flag, _vp(out), pb.n_out,
                _vp(pb.curve_base_d), _vp(pb.curve_count_d), _vp(summary), ctypes.byref(handle),
            )
            self.ctx.check(st, "graph capture")
        self.handle = handle

    def launch(self) -> None:
        import torch

        stream = ctypes.c_void_p(torch.cuda.current_stream(self.out.device).cuda_stream)
        with self.ctx.lock:
            self.ctx.check(self.ctx.lib.cpb_graph_launch(self.ctx.handle, self.handle, stream),
                           "graph launch")

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.ctx.lib.cpb_graph_destroy(self.handle)
        self.handle = None

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


def capture(pb: PackedBatch, out=None, summary=None) -> BatchGraph:
    """Capture ``pb``'s search (and summaries) into a replayable :class:`BatchGraph`."""
    import torch

    device = _device()
    if out is None:
        out = torch.empty((_lib.N_OUTPUTS, max(1, pb.n_out)), dtype=torch.float64, device=device)
    if summary is None:
        summary = torch.empty((len(pb.plans), 4), dtype=torch.float64, device=device)
    return BatchGraph(pb, out, summary)


def finish(pb: PackedBatch, out, summary, *, spectra: bool = False) -> BatchResult:
    """Bring the per-curve summaries (and, on request, the spectra) to the host."""
    s = summary.cpu().numpy()
    best_index = s[:, 2].astype(np.int64)
    best_period = np.array(
        [pb.plans[c].period_at(int(best_index[c])) if pb.curve_count[c] else np.nan
         for c in range(len(pb.plans))]
    )
    result = BatchResult(
        best_period=best_period,
        best_power=s[:, 3].copy(),
        best_index=best_index,
        mean=s[:, 0].copy(),
        std=s[:, 1].copy(),
        n_periods=pb.curve_count.copy(),
        plans=pb.plans,
    )
    if spectra:
        host = out.cpu().numpy()
        result.spectra = []
        for c in range(len(pb.plans)):
            sl = slice(int(pb.curve_base[c]), int(pb.curve_base[c] + pb.curve_count[c]))
            result.spectra.append(
                {name: host[i, sl].copy() for i, name in enumerate(core._OUTPUT_NAMES)}
            )
    return result


def bls_batch(
    lcs: Sequence[tuple[np.ndarray, np.ndarray, np.ndarray]],
    grids: Sequence[LinspaceGrid | GridSpec | None] | LinspaceGrid | GridSpec | None = None,
    settings: BLSSettings | None = None,
    *,
    spectra: bool = False,
) -> BatchResult:
    """BLS of many light curves ``(t, flux, dy)`` (flux domain) in one launch sequence.

    ``grids`` is one grid for every curve or one per curve (``None`` = default grid).
    """
    settings = settings or BLSSettings()
    if grids is None or isinstance(grids, (LinspaceGrid, GridSpec)):
        grids = [grids] * len(lcs)
    if len(grids) != len(lcs):
        raise ValueError("grids must be one grid or one per light curve")
    if not lcs:
        empty = np.zeros(0)
        return BatchResult(empty, empty, empty.astype(np.int64), empty, empty,
                           empty.astype(np.int64), [], [] if spectra else None)
    pb = pack(lcs, grids, settings)
    out, summary = run_packed(pb)
    return finish(pb, out, summary, spectra=spectra)


__all__ = ["BatchGraph", "BatchResult", "CURVE_DTYPE", "CurvePlan", "CurveSet", "LinspaceGrid",
           "PackedBatch", "PeakCandidates", "SEG_DTYPE", "bls_batch", "capture", "curve_array",
           "finish", "pack", "pack_curves", "pack_plans", "peak_candidates",
           "periods_of_columns", "plan_grid", "plan_grids", "run_packed", "sub_curves",
           "window_batch", "window_plan"]
