"""Host logic of the learned proposer: coarse grid, cells, windows and their ranges."""

from __future__ import annotations

import numpy as np
import pytest

from cuperiod_amd.batch import engine
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.learned import proposer as lp
from cuperiod_amd.learned.network import L, init_parameters, load_parameters, save_parameters
from cuperiod_amd.methods.bls import _log_segments


def test_coarse_grid_and_cells():
    f = lp.coarse_frequencies(0.002, 2.0)
    assert f.size == L and abs(f[0] - 0.002) < 1e-15 and abs(f[-1] - 2.0) < 1e-12
    assert np.all(np.diff(np.log(f)) > 0)
    x = lp.coarse_position(f, np.full(L, 0.002), np.full(L, 2.0))
    np.testing.assert_array_equal(x, np.arange(L))
    mid = np.sqrt(f[100] * f[101]) * 0.999
    assert lp.coarse_position(np.array([mid]), np.array([0.002]), np.array([2.0]))[0] == 100
    assert lp.coarse_position(np.array([5.0, 1e-5]), np.array([0.002] * 2),
                              np.array([2.0] * 2)).tolist() == [L - 1, 0]


def test_coarse_plan_uses_the_plain_segments():
    s = BLSSettings(n_durations=6)
    grid = engine.LinspaceGrid(1 / 700.0, 2.0, 200_000)
    plan = lp.coarse_plan(0, grid, s)
    periods = plan.periods()
    assert periods.size == L and np.all(np.diff(periods) > 0)
    full = engine._linspace_plan(0, grid, s)
    full_durs = {round(float(sg.durations.min()), 12) for sg in full.segments}
    for sg in plan.segments:
        assert round(float(sg.durations.min()), 12) in full_durs
    ref = _log_segments(periods, s)
    assert [p.size for p, _ in ref] == [sg.n_periods for sg in plan.segments]


def _prep_for(grids, fine_per_window=2000):
    f_lo, dlog, half = lp.window_geometry(grids, fine_per_window)
    return lp.PreparedGrids(cs=None, grids=list(grids), full=None, coarse=None,
                            geometry_host=(f_lo, dlog, half))


@pytest.mark.parametrize("fine_per_window", [2000, 20])
def test_refine_ranges_hold_every_claimed_cells_target(fine_per_window):
    # the refined index range must hold every grid frequency within the half-width of
    # every claimed cell, and every frequency whose nearest cell is claimed; at 20 fine
    # frequencies per window the half-width is far below half a coarse cell
    grids = [engine.LinspaceGrid(0.001, 2.0, 200_000), engine.LinspaceGrid(0.05, 2.0, 50_000)]
    prep = _prep_for(grids, fine_per_window)
    f_lo, dlog, half = prep.geometry_host
    spans = [(0, 0, 0), (0, 1500, 1503), (0, L - 3, L - 1), (1, 0, 2), (1, 2000, 2001)]
    curve = np.array([c for c, _, _ in spans])
    lo = np.array([a for _, a, _ in spans])
    hi = np.array([b for _, _, b in spans])
    k0, k1 = lp.refine_ranges(prep, lo, hi, curve)
    for j, (c, a, b) in enumerate(spans):
        g = grids[c]
        freq = g.frequencies()
        cells = f_lo[c] * np.exp(np.arange(a, b + 1) * dlog[c])
        need = np.zeros(freq.size, bool)
        for fc in cells:
            need |= np.abs(freq - fc) <= half[c]
        nearest = np.rint(np.log(freq / f_lo[c]) / dlog[c])
        need |= (nearest >= a) & (nearest <= b)
        i = np.flatnonzero(need)
        k = g.num - 1 - i  # ascending-period index
        assert k.size == 0 or (k0[j] <= k.min() and k1[j] > k.max()), (j, k0[j], k1[j])
        assert k1[j] - k0[j] <= k.size + 8


def test_merge_ranges():
    curve = np.array([1, 0, 1, 0, 1, 2])
    k0 = np.array([50, 10, 40, 30, 90, 5])
    k1 = np.array([60, 20, 55, 35, 95, 6])
    c, a, b = lp._merge_ranges(curve, k0, k1)
    assert list(zip(c.tolist(), a.tolist(), b.tolist())) == [
        (0, 10, 20), (0, 30, 35), (1, 40, 60), (1, 90, 95), (2, 5, 6)]
    c, a, b = lp._merge_ranges(np.array([3, 3]), np.array([0, 10]), np.array([10, 12]))
    assert list(zip(c.tolist(), a.tolist(), b.tolist())) == [(3, 0, 12)]


def _no_device(monkeypatch):
    monkeypatch.setattr(engine, "_device", lambda: None)
    monkeypatch.setattr(engine, "_upload",
                        lambda a, dtype, device: np.ascontiguousarray(a, dtype=dtype))


def test_window_batch_cuts_the_full_descriptors(monkeypatch):
    # windows cut from the packed descriptor table search the same periods, with the
    # same binning and reference segment, as window_plan + pack_plans
    _no_device(monkeypatch)
    from cuperiod_amd.workloads import survey_curves

    s = BLSSettings(n_durations=6)
    curves, grids, _ = survey_curves(3, seed=4, n_freq=60_000)
    grids[1] = grids[1].as_gridspec()  # one explicit grid
    cs = engine.pack_curves(curves)
    plans = engine.plan_grids(cs, grids, s)
    full = engine.pack_plans(cs, plans, s)
    wins = {0: [(0, 700), (12_000, 13_500)], 1: [(100, 200), (59_000, 60_000)],
            2: [(30_000, 31_000)]}
    curve = np.array([c for c in wins for _ in wins[c]])
    k0 = np.array([a for c in wins for a, _ in wins[c]])
    k1 = np.array([b for c in wins for _, b in wins[c]])
    wb = engine.window_batch(full, curve, k0, k1)
    got = engine.periods_of_columns(wb, np.arange(wb.n_out))
    want = np.concatenate([plans[c].periods()[a:b] for c in wins for a, b in wins[c]])
    np.testing.assert_array_equal(got, want)
    # every new column maps back to the same period of the full batch
    tab = wb.segment_table[:wb.n_segments]
    r = np.searchsorted(tab["out_offset"], np.arange(wb.n_out), side="right") - 1
    src = wb.source_col[r] + (np.arange(wb.n_out) - tab["out_offset"][r])
    np.testing.assert_array_equal(engine.periods_of_columns(full, src), got)
    # binning, box widths and reference segment follow the period's full-batch piece
    ftab = full.segment_table[:full.n_segments]
    fr = np.searchsorted(ftab["out_offset"], src, side="right") - 1
    for f in ("bin_duration", "dur_offset", "n_dur", "curve", "group", "mode", "width"):
        np.testing.assert_array_equal(tab[f][r], ftab[f][fr], err_msg=f)
    assert wb.curve_count.tolist() == [2200, 1100, 1000]


def test_time_chunks(monkeypatch):
    _no_device(monkeypatch)
    rng = np.random.default_rng(1)
    t = np.sort(rng.uniform(0, 300, 3000))
    t = t[(t < 100) | (t > 160)]  # a gap
    lcs = [(t, np.ones(t.size), np.ones(t.size)), (np.sort(rng.uniform(0, 5, 40)),
                                                    np.ones(40), np.ones(40))]
    cs = engine.pack_curves(lcs)
    parent, start, count = lp.time_chunks(cs, np.array([40.0, 40.0]))
    for p_, a, n in zip(parent, start, count):
        tt = lcs[p_][0][a:a + n]
        assert tt[-1] - tt[0] <= 40.0 and n >= lp.MIN_CHUNK_POINTS
    first = parent == 0
    assert count[first].sum() == t.size and np.all(np.diff(start[first]) > 0)
    assert parent.tolist().count(1) == 1


def test_window_plan_keeps_segment_binning():
    s = BLSSettings(n_durations=6)
    grid = engine.LinspaceGrid(1 / 900.0, 2.0, 50_000)
    full = engine._linspace_plan(0, grid, s)
    per = full.periods()
    ranges = [(10, 40), (100, 5000), (49_990, 50_000)]
    wp = engine.window_plan(full, ranges)
    want = np.concatenate([per[a:b] for a, b in ranges])
    np.testing.assert_array_equal(wp.periods(), want)
    assert wp.n_periods == want.size
    for sg in wp.segments:
        src = [f for f in full.segments if f.lin_k0 <= sg.lin_k0 < f.lin_k0 + f.n_periods][0]
        assert sg.p_max == src.p_max
        np.testing.assert_array_equal(sg.durations, src.durations)
    for k in (0, 29, 30, 4929, want.size - 1):
        assert wp.period_at(k) == want[k]


def test_select_windows_by_mass():
    prop = lp.LearnedProposer.__new__(lp.LearnedProposer)
    prop.window_mass = 0.9
    centre = np.array([[5, 9, 100, -1], [7, 8, 9, 10], [3, 4, -1, -1]])
    mass = np.array([[0.45, 0.5, 0.01, 0.0], [0.2, 0.2, 0.2, 0.2], [0.95, 0.04, 0, 0]],
                    dtype=np.float32)
    keep, conf = prop.select(centre, mass)
    assert keep[0].tolist() == [True, True, False, False] and abs(conf[0] - 0.95) < 1e-6
    assert keep[1].tolist() == [True] * 4 and abs(conf[1] - 0.8) < 1e-6
    assert keep[2].tolist() == [True, False, False, False] and abs(conf[2] - 0.95) < 1e-6


def test_parameters_round_trip(tmp_path):
    p = init_parameters(3)
    save_parameters(p, tmp_path / "w.npz")
    q = load_parameters(tmp_path / "w.npz")
    assert p.keys() == q.keys()
    for k in p:
        np.testing.assert_array_equal(p[k], q[k])


def test_label_ranges_cover_the_refinement_window():
    grids = [engine.LinspaceGrid(1 / 500.0, 2.0, 200_000), engine.LinspaceGrid(0.05, 2.0, 200_000)]
    f_lo, dlog, half = lp.window_geometry(grids, 2000)
    np.testing.assert_allclose(half, 1000 * np.array([(2.0 - 1 / 500.0), 1.95]) / 199_999)
    freq = np.array([0.01, 1.9])
    r = lp.label_ranges(freq, f_lo, dlog, half)
    for b in range(2):
        cells = f_lo[b] * np.exp(np.arange(L) * dlog[b])
        inside = np.flatnonzero(np.abs(cells - freq[b]) <= half[b])
        if inside.size:
            assert r[b, 0] <= inside[0] and r[b, 1] >= inside[-1]
            assert r[b, 1] - r[b, 0] <= inside.size + 1
        near = int(np.argmin(np.abs(np.log(cells / freq[b]))))
        assert r[b, 0] <= near <= r[b, 1]
