"""Host logic of the learned proposer: coarse grid, cells, windows and their ranges."""

from __future__ import annotations

import numpy as np

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


def test_window_ranges_merge_and_clip():
    grid = engine.LinspaceGrid(0.01, 2.0, 200_000)
    f = lp.coarse_frequencies(0.01, 2.0)
    step = (grid.stop - grid.start) / (grid.num - 1)
    r = lp.window_ranges(np.array([L - 1]), grid, 0.01, 2.0, 1000)
    assert r == [(0, 1000)]  # highest frequency = shortest period, clipped at index 0
    r = lp.window_ranges(np.array([0]), grid, 0.01, 2.0, 1000)
    assert r == [(grid.num - 1 - 1000, grid.num)]
    c = np.array([2000, 2001])  # neighbouring cells: overlapping windows merge
    r = lp.window_ranges(c, grid, 0.01, 2.0, 1000)
    assert len(r) == 1 and r[0][1] - r[0][0] > 2000
    k = grid.num - 1 - int(np.rint((f[2000] - grid.start) / step))
    assert r[0][0] <= k - 1000 and r[0][1] >= k + 1000


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
    centre = np.array([[5, 9, 100, -1], [7, 8, 9, 10]])
    mass = np.array([[0.5, 0.45, 0.01, 0.0], [0.2, 0.2, 0.2, 0.2]], dtype=np.float32)
    kept, conf = prop.select(centre, mass)
    assert kept[0].tolist() == [5, 9] and abs(conf[0] - 0.95) < 1e-6
    assert kept[1].tolist() == [7, 8, 9, 10] and abs(conf[1] - 0.8) < 1e-6


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
