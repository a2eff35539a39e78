"""Host-side logic (no GPU): grids, segmentation, settings, results and peak
selection, error mapping, and the C ABI's exported symbols."""

from __future__ import annotations

import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest

import cuperiod_amd as cup
from cuperiod_amd import _lib
from cuperiod_amd.core.grid import GridSpec, check_grid_size, linspace_frequencies
from cuperiod_amd.core.peaks import (
    is_alias_related,
    local_maxima,
    peak_preserving_downsample,
    select_diverse_peaks,
    select_top_peaks,
)
from cuperiod_amd.methods import _bls_core as core
from cuperiod_amd.methods.bls import (
    BLSMethod,
    _grid_segments,
    _log_segments,
    _segment_durations,
    _segment_periods,
)
from synth import synthetic_eclipser

ROOT = Path(__file__).resolve().parents[1]


def _flux_lc(t, flux, err):
    return cup.LightCurve.from_arrays(t, flux, err, domain=cup.Domain.FLUX)


# --- C ABI ------------------------------------------------------------------------


def _header_functions(header: str = "cuperiod_bls.h", prefix: str = "cpb") -> list[str]:
    text = (ROOT / "include" / header).read_text()
    return re.findall(rf"^\s*(?:int|void|const char\*)\s+({prefix}_\w+)\(", text, re.M)


def test_header_and_binding_declare_the_same_symbols():
    assert tuple(_header_functions()) == _lib.EXPORTED_SYMBOLS
    assert tuple(_header_functions("cuperiod_proposer.h", "cpp")) == _lib.PROPOSER_SYMBOLS


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in _header_functions() + _header_functions("cuperiod_proposer.h", "cpp"):
        assert hasattr(lib, name), name


def test_library_host_entry_points():
    lib = _lib.load_library()
    assert lib.cpb_version() >= 100
    assert lib.cpb_status_string(0) == b"ok"
    assert lib.cpb_status_string(3) == b"resource limit exceeded"
    out = (ctypes.c_double * 1001)()
    assert lib.cpb_linspace_grid(0.0005, 2.0, 1001, out) == 0
    np.testing.assert_array_equal(np.frombuffer(out, dtype=np.float64),
                                  np.linspace(0.0005, 2.0, 1001))


def test_segment_struct_layout_matches_header():
    assert ctypes.sizeof(_lib.Segment) == 96
    assert _lib.Segment.lin_start.offset == 56
    assert ctypes.sizeof(_lib.Curve) == 48
    assert _lib.Curve.sum_ivar.offset == 40


def test_struct_layouts_match_the_c_compiler(tmp_path):
    # sizes and offsets as a C compiler lays out the header's structs
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "cuperiod_bls.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu\\n", sizeof(cpb_segment), sizeof(cpb_curve),\n'
        "         sizeof(cpb_band), sizeof(cpb_star), sizeof(cpb_tls_curve));\n"
        '  printf("%zu %zu %zu %zu\\n", offsetof(cpb_segment, lin_start),\n'
        "         offsetof(cpb_star, t_min), offsetof(cpb_tls_curve, w_total),\n"
        "         offsetof(cpb_curve, sum_y));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    sizes, offsets = subprocess.run([str(exe)], check=True, capture_output=True,
                                    text=True).stdout.splitlines()
    assert [int(v) for v in sizes.split()] == [
        ctypes.sizeof(_lib.Segment), ctypes.sizeof(_lib.Curve), ctypes.sizeof(_lib.Band),
        ctypes.sizeof(_lib.Star), ctypes.sizeof(_lib.TlsCurve)]
    assert [int(v) for v in offsets.split()] == [
        _lib.Segment.lin_start.offset, _lib.Star.t_min.offset, _lib.TlsCurve.w_total.offset,
        _lib.Curve.sum_y.offset]


@pytest.mark.parametrize("start,stop,num", [(1 / 50, 2.0, 200000), (0.001, 5.0, 1000000),
                                             (1 / 1337.5, 2.0, 777), (0.3, 0.3000001, 5)])
def test_linspace_recipe_is_numpys(start, stop, num):
    np.testing.assert_array_equal(linspace_frequencies(start, stop, num),
                                  np.linspace(start, stop, num))


# --- grids ------------------------------------------------------------------------


def test_gridspec_views_and_validation():
    g = GridSpec(kind="frequency", values=np.linspace(0.1, 1.0, 10), uniform=True)
    np.testing.assert_allclose(g.period, 1.0 / np.linspace(0.1, 1.0, 10)[::-1])
    assert g.uniform_frequency_params()[2] == 10
    with pytest.raises(ValueError, match="ascending"):
        GridSpec(kind="period", values=np.array([3.0, 2.0, 1.0]))
    with pytest.raises(ValueError, match="> 0"):
        GridSpec(kind="period", values=np.array([0.0, 1.0]))
    with pytest.raises(ValueError, match="finite"):
        GridSpec(kind="frequency", values=np.array([0.1, np.inf]))


def test_grid_size_check_refuses_before_allocation():
    with pytest.raises(cup.GridTooLargeError):
        check_grid_size(10**12, context="test")


def test_default_grid_has_no_overshooting_period():
    # tests/test_bls.py:427-437 of the reference.
    t = 2458000.0 + np.linspace(0.0, 115.35, 400)
    rng = np.random.default_rng(0)
    lc = _flux_lc(t, 1.0 + 0.01 * rng.standard_normal(t.size), np.full(t.size, 0.01))
    settings = cup.BLSSettings()
    grid = BLSMethod().default_grid(lc, settings)
    assert float(grid.period.min()) > settings.min_period_days
    assert bool(np.all(np.diff(grid.period) > 0.0))
    assert sum(grid.meta["segment_sizes"]) == grid.size


def test_default_grid_is_size_checked():
    # tests/test_bls.py:440-446: one t=0 row inflates the baseline.
    t, flux, err = synthetic_eclipser(n=200)
    t[0] = 0.0
    with pytest.raises(cup.GridTooLargeError, match="BLS default grid"):
        BLSMethod().default_grid(_flux_lc(t, flux, err), cup.BLSSettings())


def test_default_segments_closed_form():
    settings = cup.BLSSettings()
    baseline = 60.0
    parts = _segment_periods(baseline, settings)
    df = settings.grid_duration_frac / (settings.grid_oversample * baseline)
    p_lo = settings.min_period_days
    for lo, periods in parts:
        assert lo == pytest.approx(p_lo)
        p_hi = min(lo * 2, baseline / 2)
        freq = 1.0 / periods[::-1]
        assert freq[0] == pytest.approx(1.0 / p_hi)
        assert np.all(freq < 1.0 / lo)
        np.testing.assert_allclose(np.diff(freq), df, rtol=1e-6)
        p_lo = p_hi
    d = _segment_durations(1.0, settings)
    np.testing.assert_allclose(d, np.geomspace(0.01, 0.12, 8))
    assert _segment_durations(0.05, settings).tolist() == [0.05 * 0.12]


def test_custom_grid_log_segments_keep_every_period():
    settings = cup.BLSSettings()
    periods = np.geomspace(1.0, 100.0, 4000)
    segs = _log_segments(periods, settings)
    np.testing.assert_array_equal(np.concatenate([p for p, _ in segs]), periods)
    for k, (p, d) in enumerate(segs):
        assert p.min() >= 1.0 * 2**k * (1 - 1e-12)
        assert p.max() < 1.0 * 2 ** (k + 1)
        np.testing.assert_allclose(d, _segment_durations(2.0**k, settings))


def test_grid_segments_use_recorded_meta():
    t, flux, err = synthetic_eclipser(span=80.0)
    lc = _flux_lc(t, flux, err)
    settings = cup.BLSSettings(min_period_days=1.0, max_period_days=4.0)
    grid = BLSMethod().default_grid(lc, settings)
    segs = _grid_segments(grid, settings)
    assert [p.size for p, _ in segs] == list(grid.meta["segment_sizes"])


def test_frequency_grid_through_zero_is_refused():
    t, flux, err = synthetic_eclipser()
    grid = GridSpec(kind="frequency", values=np.linspace(0.0, 0.5, 200))
    with pytest.raises(ValueError, match="finite"):
        BLSMethod().power(grid, _flux_lc(t, flux, err), cup.BLSSettings(), "hip")


def test_empty_grid_raises():
    t, flux, err = synthetic_eclipser()
    with pytest.raises(cup.InsufficientDataError, match="empty trial grid"):
        BLSMethod().power(GridSpec(kind="period", values=np.zeros(0)),
                          _flux_lc(t, flux, err), cup.BLSSettings(), "hip")


def test_too_few_points_raise():
    t, flux, err = synthetic_eclipser(n=10)
    grid = GridSpec(kind="period", values=np.linspace(1, 2, 10))
    with pytest.raises(cup.InsufficientDataError, match="min_detections"):
        BLSMethod().power(grid, _flux_lc(t, flux, err), cup.BLSSettings(), "hip")


# --- host inputs / spec -------------------------------------------------------------


def test_host_inputs_and_spec():
    t, y, ivar = core._host_inputs([3.0, 1.0, 2.0], [5.0, 1.0, 2.0], [0.5, 1.0, 2.0])
    np.testing.assert_array_equal(y, [3.0, -1.0, 0.0])
    np.testing.assert_array_equal(ivar, [4.0, 1.0, 0.25])
    spec = core._search_spec(np.array([1.0, 2.0]), np.array([0.2, 0.25, 0.45]), 2)
    assert spec.bin_duration == 0.1
    assert spec.dur_bins == [2, 3, 5]
    assert spec.width == math.ceil(2.0 / 0.1) + 2 + 1
    assert core._ivar_floor(100, 0.0) == np.finfo(np.float64).tiny
    assert core._ivar_floor(100, 2.0) == 4.0 * 100 * np.finfo(np.float64).eps * 2.0


def test_unknown_backend_and_objective():
    with pytest.raises(ValueError, match="objective"):
        core.bls_power(np.arange(5.0), np.ones(5), np.ones(5), np.array([1.0]),
                       np.array([0.1]), 2, objective="chi2")
    with pytest.raises(ValueError, match="unknown backend"):
        core.bls_power(np.arange(5.0), np.ones(5), np.ones(5), np.array([1.0]),
                       np.array([0.1]), 2, backend="numba")
    out = core.bls_power(np.arange(5.0), np.ones(5), np.ones(5), np.zeros(0),
                         np.array([0.1]), 2)
    assert out.power.size == 0


# --- backend resolution -------------------------------------------------------------


def test_cpu_backends_are_refused():
    m = cup.get_method("bls")
    for name in ("cpu", "numba", "numpy", "astropy", "cupy", "torch:cpu"):
        with pytest.raises(cup.BackendUnavailableError):
            m.resolve_backend(name)
    with pytest.raises(cup.BackendUnavailableError, match="unknown backend"):
        m.resolve_backend("fortran")


def test_unknown_method():
    with pytest.raises(cup.UnknownMethodError):
        cup.get_method("GLS")
    assert cup.get_method("b-l-s").name == "BLS"


# --- settings ---------------------------------------------------------------------


def test_settings_defaults_bounds_and_environment(monkeypatch):
    s = cup.BLSSettings()
    assert (s.min_period_days, s.max_period_days, s.n_durations, s.bins_per_duration) == (
        0.2, 100.0, 8, 10)
    assert s.alias_freq_tolerance == 0.0035
    with pytest.raises(ValueError):
        cup.BLSSettings(min_period_days=5.0, max_period_days=1.0)
    with pytest.raises(ValueError):
        cup.BLSSettings(duration_max_frac=1.5)
    with pytest.raises(ValueError):
        cup.BLSSettings(no_such_field=1)
    monkeypatch.setenv("CUPERIOD_BLS_N_DURATIONS", "6")
    monkeypatch.setenv("CUPERIOD_BLS_OBJECTIVE", "likelihood")
    s = cup.BLSSettings()
    assert s.n_durations == 6 and s.objective == "likelihood"
    assert cup.BLSSettings(n_durations=4).n_durations == 4


# --- results and peaks ------------------------------------------------------------


def test_from_spectrum_sorts_by_frequency_stably():
    freq = np.array([3.0, 1.0, 2.0, 1.0])
    power = np.array([0.3, 0.1, 0.2, 0.15])
    pg = cup.Periodogram.from_spectrum(method="BLS", backend="hip", frequency=freq,
                                       power=power, objective_sense="max", n_samples=5,
                                       baseline=10.0, extras={"x": np.arange(4.0)})
    np.testing.assert_array_equal(pg.frequency, [1.0, 1.0, 2.0, 3.0])
    np.testing.assert_array_equal(pg.extras["x"], [1.0, 3.0, 2.0, 0.0])


def _brute_candidates(s):
    n = s.size
    out = [i for i in range(1, n - 1) if s[i] > s[i - 1] and s[i] >= s[i + 1]]
    if n >= 2 and s[0] >= s[1]:
        out.append(0)
    if n >= 2 and s[-1] > s[-2]:
        out.append(n - 1)
    return out


def test_best_periods_against_brute_force_greedy():
    rng = np.random.default_rng(3)
    freq = np.sort(rng.uniform(0.05, 3.0, 600))
    power = rng.random(600)
    pg = cup.Periodogram.from_spectrum(method="BLS", backend="hip", frequency=freq,
                                       power=power, objective_sense="max", n_samples=50,
                                       baseline=40.0)
    cands = _brute_candidates(pg.power)
    assert sorted(pg._peak_candidates(pg.power).tolist()) == sorted(cands)
    sep = 3.0 / 40.0
    chosen = []
    for i in sorted(cands, key=lambda i: -pg.power[i]):
        if all(abs(pg.frequency[i] - pg.frequency[j]) >= sep for j in chosen):
            chosen.append(i)
        if len(chosen) == 5:
            break
    assert [p.frequency for p in pg.best_periods(5)] == [pg.frequency[i] for i in chosen]
    assert pg.best_period() == pg.period[chosen[0]]


def test_alias_relations():
    assert is_alias_related(2.0, 1.0, 1e-3, 8)          # harmonic
    assert is_alias_related(0.5, 1.0, 1e-3, 8)          # sub-harmonic
    assert is_alias_related(1.2 + 1.0027379, 1.2, 1e-3, 8)  # sidereal alias
    assert is_alias_related(1.2 + 1 / 365.25, 1.2, 1e-3, 8)  # yearly sidelobe
    assert not is_alias_related(1.37, 1.0, 1e-3, 8)
    freq = np.array([0.5, 1.0, 1.37, 2.0])
    score = np.array([3.0, 4.0, 1.0, 2.0])
    cand = np.arange(4)
    assert select_diverse_peaks(freq, score, cand, 4, 1e-3, 8).tolist() == [1, 2]
    assert select_top_peaks(freq, score, cand, 2, 0.1).tolist() == [1, 0]


def test_local_maxima_and_downsample():
    s = np.array([0.0, 2.0, 2.0, 1.0, np.nan, 3.0, 0.0])
    assert local_maxima(s).tolist() == [1, 5]
    f = np.arange(10000.0)
    p = np.random.default_rng(0).random(10000)
    p[1234] = 5.0
    fd, pd = peak_preserving_downsample(f, p, 100)
    assert pd.max() == np.float32(5.0) and fd[np.argmax(pd)] == np.float32(1234.0)


def test_periods_at_equals_the_materialised_grid():
    """CurvePlan.periods_at gives the periods() array's values bit for bit at any
    indices (unordered, repeated, segment edges) of linspace, default and explicit plans
    and of a plan mixing explicit and parametric segments."""
    from cuperiod_amd.batch.engine import (
        CurvePlan, LinspaceGrid, _default_plan, _explicit_plan, _linspace_plan)
    from cuperiod_amd.core.grid import GridSpec

    settings = cup.BLSSettings(n_durations=6)
    rng = np.random.default_rng(8)
    t = np.sort(rng.uniform(0.0, 300.0, 500))
    plans = [_linspace_plan(0, LinspaceGrid(1.0 / 700.0, 2.0, 30_000), settings),
             _default_plan(1, float(t[-1] - t[0]), t, settings),
             _explicit_plan(2, GridSpec(kind="period",
                                        values=np.sort(rng.uniform(0.3, 40.0, 5_000))), settings)]
    plans.append(CurvePlan(segments=plans[2].segments + plans[0].segments,
                           n_periods=plans[2].n_periods + plans[0].n_periods))
    for plan in plans:
        every = plan.periods()
        edges = np.cumsum([0] + [sg.n_periods for sg in plan.segments])
        k = np.concatenate([rng.integers(0, plan.n_periods, 4_000), edges[:-1], edges[1:] - 1,
                            np.arange(plan.n_periods)[::-7]])
        assert np.array_equal(plan.periods_at(k).view(np.int64), every[k].view(np.int64))
    with pytest.raises(IndexError):
        plans[0].periods_at(np.array([plans[0].n_periods]))


def test_parametric_width_cuts_equal_materialised_cuts():
    # The batch planner cuts parametric (linspace) segments at the engine's span classes
    # by bisection on the period index; the cuts must equal those of the materialised
    # periods (core._width_ranges), including segments that cross several classes.
    from cuperiod_amd.batch.engine import (
        CurvePlan, LinspaceGrid, _linspace_plan, _parametric_width_ranges)

    settings = cup.BLSSettings(n_durations=6)
    rng = np.random.default_rng(2)
    sgs = []
    for c, base in enumerate(rng.uniform(50.0, 2000.0, 6)):
        plan = _linspace_plan(c, LinspaceGrid(1.0 / base, 2.0, 20_000), settings)
        sgs += [sg for sg in plan.segments if sg.mode != 0 and sg.n_periods >= 2]
    got = _parametric_width_ranges(sgs, settings.bins_per_duration)
    crossing = 0
    for sg, ranges in zip(sgs, got):
        bd = float(sg.durations.min()) / settings.bins_per_duration
        per = CurvePlan(segments=[sg], n_periods=sg.n_periods).periods_at(
            np.arange(sg.n_periods))
        want = core._width_ranges(per, bd, settings.bins_per_duration)
        ranges = [tuple(r) for r in ranges.tolist()]
        if len(want) == 1:
            assert len(ranges) == 1 and ranges[0][:2] == (0, sg.n_periods)
        else:
            assert ranges == want
            crossing += 1
    assert crossing >= 6


def test_native_spectrum_candidates_equal_the_periodogram_rule():
    """cpb_spectrum_candidates lists an engine power row's candidates (indices into the
    ascending-frequency spectrum, in the order the selection sees them) and scores as
    Periodogram._peak_candidates does on the row reversed with non-finite powers read as
    0: flat tops, plateaus at both ends, NaN/inf, 0-3 samples and many exact ties."""
    import ctypes

    from cuperiod_amd import _lib
    from cuperiod_amd.core.result import Periodogram

    lib = _lib.load_library()
    rng = np.random.default_rng(5)
    rows = [np.zeros(0), np.array([3.0]), np.array([1.0, 2.0]), np.array([2.0, 1.0]),
            np.array([1.0, 1.0]), np.array([1.0, 2.0, 3.0]), np.array([3.0, 2.0, 1.0]),
            np.array([np.nan, np.inf, -np.inf]), np.full(7, 4.0), np.array([1.0, 3.0, 3.0, 1.0]),
            np.tile([0.0, 1.0], 500), np.append(np.tile([1.0, 0.0], 500), 1.0)]
    for trial in range(400):
        n = int(rng.integers(1, 3000))
        row = np.round(rng.gamma(2.0, 1.0, n) * int(rng.integers(1, 30))) if trial % 2 else \
            rng.normal(size=n)
        bad = rng.random(n) < 0.02
        row[bad] = rng.choice([np.nan, np.inf, -np.inf], int(bad.sum()))
        rows.append(row)
    for row in rows:
        n = row.size
        score = np.nan_to_num(row[::-1], nan=0.0, posinf=0.0, neginf=0.0)
        freq = np.linspace(0.1, 2.0, n)
        pg = Periodogram(method="BLS", backend="hip", frequency=freq, period=1.0 / freq,
                         power=score, objective_sense="max", n_samples=10, baseline=50.0)
        want = pg._peak_candidates(score)
        r = np.ascontiguousarray(row, dtype=np.float64)
        index = np.empty(n // 2 + 2, dtype=np.int64)
        got = np.empty(n // 2 + 2, dtype=np.float64)
        count = ctypes.c_int64()
        assert lib.cpb_spectrum_candidates(r.ctypes.data, n, index.ctypes.data,
                                           got.ctypes.data, ctypes.byref(count)) == 0
        m = count.value
        assert np.array_equal(index[:m], want), row[:12]
        assert np.array_equal(got[:m], score[want])
        # the order numpy's argsort gives the candidates' scores is the same
        assert np.array_equal(index[:m][np.argsort(got[:m])[::-1]],
                              want[np.argsort(score[want])[::-1]])


def test_native_alias_diverse_selection_equals_numpy_rule():
    """The library's host-side greedy (cpb_select_diverse_peaks) chooses the same peaks,
    in the same order, and visits as many candidates as the numpy rule, on spectra
    built around harmonics, sub-harmonics and day/year aliases of a base frequency."""
    from cuperiod_amd.core import peaks

    assert peaks._native_library() is not None
    rng = np.random.default_rng(11)
    for trial in range(600):
        n = int(rng.integers(5, 400))
        base = rng.uniform(0.05, 1.5)
        f = np.abs(np.sort(np.concatenate([
            rng.uniform(0.001, 5, n // 2),
            base * rng.integers(1, 6, n - n // 2) + rng.normal(0, 2e-3, n - n // 2),
            base / rng.integers(1, 6, 5),
            base + np.array([1.0027379, 1.0, 2.0, 1 / 365.25, -1.0])]))) + 1e-4
        score = rng.normal(size=f.size)
        if trial % 4 == 0:
            score = np.round(score, 1)
        cand = rng.permutation(f.size)[: max(1, f.size // 2)]
        tol = float(rng.choice([0.0035, 0.01, 0.05, 0.2]))
        args = (f, score, cand, int(rng.integers(0, 25)), tol, int(rng.integers(1, 9)))
        kw = {"min_separation": None if trial % 2 else float(rng.uniform(0, 0.05))}
        seen_native, seen_numpy = [], []
        native = peaks.select_diverse_peaks(*args, examined=seen_native, **kw)
        peaks._NATIVE[:] = [None]
        try:
            rule = peaks.select_diverse_peaks(*args, examined=seen_numpy, **kw)
        finally:
            peaks._NATIVE[:] = []
        assert np.array_equal(native, rule) and seen_native == seen_numpy, trial
