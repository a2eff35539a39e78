"""Host logic of ``batch_periodograms`` on the CPU.

The device search is replaced by the oracle (the spectra a correct engine returns), so
these tests check what the batch runner adds around it: per-curve checks and errors, input
order, chunking, rows, sinks, resume and the directory manifest
(src/cuperiod/batch/runner.py:546-1032, src/cuperiod/batch/io.py:507-667).
"""

from __future__ import annotations

import json
import math

import numpy as np
import pytest

import oracle
from cuperiod_amd import BLSSettings, LightCurve
from cuperiod_amd.batch import io as bio
from cuperiod_amd.batch import runner
from cuperiod_amd.core.errors import BackendUnavailableError
from cuperiod_amd.methods import base
from cuperiod_amd.methods.bls import BLSMethod, _grid_segments


def _curve(seed: int, n: int = 160, baseline: float = 24.0, period: float = 2.3):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, baseline, n))
    y = 1.0 + rng.normal(0.0, 1e-3, n)
    y[np.mod(t, period) < 0.08] -= 4e-3
    return LightCurve(t, y, np.full(n, 1e-3), domain="flux")


def _oracle_spectra(prepared, settings, grid):
    out = []
    for p in prepared:
        g = p.grid if p.grid is not None else grid
        g = g if g is not None else BLSMethod().default_grid(p.finite, settings)
        t, y = p.finite.time, p.finite.value
        err = p.finite.error if p.finite.error is not None else np.ones_like(y)
        parts = [oracle.bls_power(t, y, err, per, dur, settings.bins_per_duration,
                                  objective=settings.objective, threads=2)
                 for per, dur in _grid_segments(g, settings)]
        spectrum = {k: np.concatenate([q[k] for q in parts]) for k in oracle.OUTPUT_NAMES}
        out.append((np.concatenate([q["period"] for q in parts]), spectrum))
    return out


@pytest.fixture
def fake_engine(monkeypatch):
    calls = []

    def launch(prepared, settings, grid):
        calls.append([p.key for p in prepared])
        return prepared, settings, grid

    monkeypatch.setattr(base, "hip_ready", lambda: True)
    monkeypatch.setattr(runner, "_launch", launch)
    monkeypatch.setattr(runner, "_collect", lambda h, **kw: _oracle_spectra(*h))
    return calls


def _expected_row(key, lc, n_best=10, settings=None, methods=("BLS",)):
    settings = settings or BLSSettings()
    prep = runner._prepare(0, key, lc, settings, None)
    (periods, spectrum), = _oracle_spectra([prep], settings, None)
    pg = runner._periodogram(prep, periods, spectrum, settings)
    return bio.periodogram_to_row(key, pg, n_best=n_best, store_raw=False,
                                  extra_keys=bio.peak_extra_columns(list(methods)), source="")


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, float) and math.isnan(x):
            assert isinstance(y, float) and math.isnan(y), k
        else:
            assert x == y, k


def test_rows_errors_and_order(fake_engine):
    lcs = [("a", _curve(1)), ("b", _curve(2, n=3)), ("c", _curve(3)),
           ("d", LightCurve([0.0, 1.0, 2.0] * 5, [1.0] * 15, [0.0] * 15, domain="flux")),
           ("e", _curve(4, baseline=30.0))]
    res = runner.batch_periodograms(lcs, "BLS", chunk_size=2)
    assert res.n_inputs == 5 and res.n_done == 3 and res.n_failed == 2
    assert [k for k, _ in res.errors] == ["b", "d"]
    assert res.errors[0][1].startswith("InsufficientDataError")
    assert [r["key"] for r in res.rows] == ["a", "c", "e"]
    assert fake_engine == [["a"], ["c"], ["e"]]
    for row, (key, lc) in zip(res.rows, [lcs[0], lcs[2], lcs[4]]):
        _same(row, _expected_row(key, lc))
    row = res.rows[0]
    assert row["backend"] == "hip" and row["method"] == "BLS" and row["source"] == ""
    assert {f"{e}_10" for e in ("period", "power", "depth", "depth_snr", "duration",
                                "t0", "sde")} <= set(row)
    assert abs(row["best_period"] - 2.3) < 0.02


def test_later_method_failure_keeps_earlier_rows(fake_engine, monkeypatch):
    # The reference's chunk loop (src/cuperiod/batch/runner.py:220-249) runs a light
    # curve's methods in order and appends each row as it is made: when a later method
    # fails, the earlier methods' rows of that curve stay; the error is recorded once.
    real_prepare_tls = runner._prepare_tls

    def prepare_tls(position, key, lc, settings, grid):
        if key == "b":
            raise ValueError("TLS refuses this light curve")
        return real_prepare_tls(position, key, lc, settings, grid)

    def tls_rows(prepared, settings, cfg):
        return [{"key": p.key, "method": "TLS"} for p in prepared]

    monkeypatch.setattr(runner, "_prepare_tls", prepare_tls)
    monkeypatch.setattr(runner, "_tls_rows", tls_rows)
    lcs = [("a", _curve(1)), ("b", _curve(2)), ("c", _curve(3))]
    res = runner.batch_periodograms(lcs, ["BLS", "TLS"], chunk_size=3)
    assert [(r["key"], r["method"]) for r in res.rows] == [
        ("a", "BLS"), ("a", "TLS"), ("b", "BLS"), ("c", "BLS"), ("c", "TLS")]
    assert res.errors == [("b", "ValueError: TLS refuses this light curve")]
    _same(res.rows[2], _expected_row("b", lcs[1][1], methods=("BLS", "TLS")))


def test_earlier_method_failure_runs_no_later_method(fake_engine, monkeypatch):
    # A failure of the first method stops the curve: no row from any method.
    def tls_rows(prepared, settings, cfg):
        return [{"key": p.key, "method": "TLS"} for p in prepared]

    monkeypatch.setattr(runner, "_tls_rows", tls_rows)
    lcs = [("a", _curve(1)), ("b", _curve(2, n=3)), ("c", _curve(3))]
    res = runner.batch_periodograms(lcs, ["BLS", "TLS"], chunk_size=3)
    assert [(r["key"], r["method"]) for r in res.rows] == [
        ("a", "BLS"), ("a", "TLS"), ("c", "BLS"), ("c", "TLS")]
    assert [k for k, _ in res.errors] == ["b"]


def test_grid_rule_gives_each_curve_its_grid(fake_engine):
    # A grid rule (a callable) gives each light curve its own grid; rows equal those of
    # the same grid passed directly.
    from cuperiod_amd.core.grid import GridSpec

    def rule(lc):
        return GridSpec(kind="frequency", values=np.linspace(1.0 / lc.baseline, 2.0, 700))

    lcs = [("a", _curve(1)), ("b", _curve(3, baseline=30.0))]
    res = runner.batch_periodograms(lcs, "BLS", grid=rule)
    assert res.n_done == 2
    for row, (key, lc) in zip(res.rows, lcs):
        one = runner.batch_periodograms([(key, lc)], "BLS", grid=rule(lc.finite()))
        _same(row, one.rows[0])
    with pytest.raises(ValueError, match="BLS only"):
        runner.batch_periodograms(lcs, ["BLS", "TLS"], grid=rule)


def test_keys_from_meta_and_renamed(fake_engine):
    a = LightCurve(*(_curve(5).time, _curve(5).value, _curve(5).error), domain="flux",
                   meta={"id": "star"})
    with pytest.warns(UserWarning, match="renamed"):
        res = runner.batch_periodograms([a, a, _curve(6)], "BLS")
    assert [r["key"] for r in res.rows] == ["star", "star#2", "2"]


def test_file_sink_resume(fake_engine, tmp_path):
    lcs = [(f"k{i}", _curve(10 + i)) for i in range(4)]
    sink = tmp_path / "out.parquet"
    res = runner.batch_periodograms(lcs[:2], "BLS", sink=sink)
    assert res.rows is None and res.n_done == 2
    res = runner.batch_periodograms(lcs, "BLS", sink=sink)
    assert res.n_skipped == 2 and res.n_done == 2
    assert fake_engine == [["k0", "k1"], ["k2", "k3"]]
    rows = bio.read_rows(sink)
    assert [r["key"] for r in rows] == ["k0", "k1", "k2", "k3"]
    _same(rows[3], _expected_row("k3", lcs[3][1]))
    res = runner.batch_periodograms(lcs, "BLS", sink=sink)
    assert res.n_skipped == 4 and res.n_workers == 0
    res = runner.batch_periodograms(lcs[:1], "BLS", sink=sink, resume=False)
    assert [r["key"] for r in bio.read_rows(sink)] == ["k0"]


def test_csv_sink_escapes_and_rejects_raw(fake_engine, tmp_path):
    sink = tmp_path / "out.csv"
    runner.batch_periodograms([("=cmd", _curve(20))], "BLS", sink=sink)
    rows = bio.read_rows(sink)
    assert rows[0]["key"] == "'=cmd"
    res = runner.batch_periodograms([("=cmd", _curve(20))], "BLS", sink=sink)
    assert res.n_skipped == 1
    with pytest.raises(ValueError, match="store_raw"):
        runner.batch_periodograms([("x", _curve(21))], "BLS", sink=sink, store_raw=True)


def test_directory_sink_manifest(fake_engine, tmp_path):
    lcs = [(f"k{i}", _curve(30 + i)) for i in range(5)]
    d = tmp_path / "parts"
    res = runner.batch_periodograms(lcs[:3], "BLS", sink=d, chunk_size=2, store_raw=True)
    assert sorted(p.name for p in d.glob("part-*.parquet")) == [
        "part-00000.parquet", "part-00001.parquet"]
    manifest = json.loads((d / "_manifest.json").read_text())
    assert manifest["config"]["n_best"] == 10 and manifest["config"]["store_raw"] is True
    rows = bio.read_rows(d / "part-00000.parquet")
    assert len(rows[0]["pgram_frequency"]) == len(rows[0]["pgram_power"]) > 0
    res = runner.batch_periodograms(lcs, "BLS", sink=d, chunk_size=2, store_raw=True)
    assert res.n_skipped == 3 and res.n_done == 2
    assert (d / "part-00002.parquet").exists()
    with pytest.raises(ValueError, match="different request"):
        runner.batch_periodograms(lcs, "BLS", sink=d, n_best=3, store_raw=True)
    with pytest.raises(ValueError, match="different request"):
        runner.batch_periodograms(lcs, "BLS", sink=d, store_raw=True,
                                  settings=BLSSettings(n_durations=4))
    res = runner.batch_periodograms(lcs[:1], "BLS", sink=d, n_best=3, resume=False)
    assert sorted(p.name for p in d.glob("part-*.parquet")) == ["part-00000.parquet"]


def test_request_checks(fake_engine, tmp_path):
    lcs = [("a", _curve(40))]
    with pytest.raises(ValueError, match="device"):
        runner.batch_periodograms(lcs, device="tpu")
    with pytest.raises(NotImplementedError):
        runner.batch_periodograms(lcs, device="hybrid")
    with pytest.raises(BackendUnavailableError):
        runner.batch_periodograms(lcs, device="cpu")
    with pytest.raises(ValueError, match="runs on the CPU"):
        runner.batch_periodograms(lcs, backend="numba")
    with pytest.raises(ValueError, match="unsupported sink"):
        runner.batch_periodograms(lcs, sink=tmp_path / "x.json")
    with pytest.raises(TypeError):
        runner.batch_periodograms("*.csv")
    with pytest.raises(ValueError, match="chunk_size"):
        runner.batch_periodograms(lcs, chunk_size=0)
    with pytest.warns(UserWarning, match="no inputs"):
        res = runner.batch_periodograms([], sink=tmp_path / "y.parquet")
    assert res.n_inputs == 0 and not (tmp_path / "y.parquet").exists()
    assert fake_engine == []


def test_no_device_refused_before_work(monkeypatch):
    monkeypatch.setattr(base, "hip_ready", lambda: False)
    with pytest.raises(BackendUnavailableError):
        runner.batch_periodograms([("a", _curve(50))], "BLS")


def _host_candidates(power, want):
    """Candidates of a power row (ascending period) by the host rule, in ascending
    frequency, with every candidate at or above the want-th largest score."""
    s = np.nan_to_num(power[::-1], nan=0.0, posinf=0.0, neginf=0.0)
    n = s.size
    x = []
    for i in range(n):
        if n == 1:
            x.append(i)
        elif i == 0:
            if s[0] >= s[1]:
                x.append(i)
        elif i == n - 1:
            if s[i] > s[i - 1]:
                x.append(i)
        elif s[i] > s[i - 1] and s[i] >= s[i + 1]:
            x.append(i)
    x = np.array(x, dtype=np.int64)
    if x.size > want:
        thr = np.sort(s[x])[::-1][want - 1]
        keep = x[s[x] >= thr]
    else:
        keep = x
    return keep, x.size


def _as_candidates(periods, spectrum, want, cap=None):
    keep, total = _host_candidates(spectrum["power"], want)
    n = periods.size
    k = n - 1 - keep
    p = np.nan_to_num(spectrum["power"], nan=0.0, posinf=0.0, neginf=0.0)
    values = np.stack([p[k], spectrum["depth"][k], spectrum["depth_snr"][k],
                       spectrum["duration"][k], spectrum["transit_time"][k]], axis=1)
    return runner._Candidates(
        periods=periods[k], values=values, total=int(total), written=int(keep.size),
        cap=cap or max(1, keep.size), mean=float(p.mean()), std=float(p.std()),
        fetch=lambda: (periods, spectrum), fetch_power=lambda: spectrum["power"].copy(),
        fetch_extras=lambda kk: np.stack([spectrum[q][kk] for q in
                                          ("depth", "depth_snr", "duration", "transit_time")],
                                         axis=1),
        periods_at=lambda k: periods[k])


@pytest.mark.parametrize("seed", [1, 3, 7, 11])
def test_rows_from_candidates_equal_full_rows(seed):
    """Rows from device-style candidates equal the full path's rows, whether the
    candidates decide the selection or the power row has to (ties, too few, truncated)."""
    settings = BLSSettings()
    lc = _curve(seed, n=300, baseline=40.0)
    prep = runner._prepare(0, "k", lc, settings, None)
    (periods, spectrum), = _oracle_spectra([prep], settings, None)
    cfg = runner._ChunkConfig(("BLS",), {"BLS": "hip"}, {"BLS": settings}, 10, False, None,
                              bio.peak_extra_columns(["BLS"]))
    full = _expected_row("k", lc)
    for want, cap_delta in ((512, 0), (3, 0), (512, -1)):
        cand = _as_candidates(periods, spectrum, want)
        if cap_delta:
            cand.cap = cand.written + cap_delta
        _same(runner._candidate_row(prep, cand, settings, cfg), full)


def test_candidates_decide_distinct_scores_without_the_power_row():
    settings = BLSSettings()
    lc = _curve(3, n=2000, baseline=40.0)
    prep = runner._prepare(0, "k", lc, settings, None)
    (periods, spectrum), = _oracle_spectra([prep], settings, None)
    cfg = runner._ChunkConfig(("BLS",), {"BLS": "hip"}, {"BLS": settings}, 10, False, None,
                              bio.peak_extra_columns(["BLS"]))
    cand = _as_candidates(periods, spectrum, 512)
    cand.fetch_power = None  # the power row must not be needed
    row = runner._candidate_row(prep, cand, settings, cfg)
    assert row is not None
    _same(row, _expected_row("k", lc))


@pytest.mark.parametrize("seed", range(12))
def test_tied_scores_decided_only_when_order_cannot_matter(seed):
    """Spectra with many exactly tied local maxima: the row built from the candidates
    (which skips the power row when no chosen peak's score is tied among the visited
    candidates) equals the row of the full power row's selection."""
    rng = np.random.default_rng(1000 + seed)
    settings = BLSSettings()
    prep = runner._prepare(0, "k", _curve(seed, n=300, baseline=40.0), settings, None)
    cfg = runner._ChunkConfig(("BLS",), {"BLS": "hip"}, {"BLS": settings}, 10, False, None,
                              bio.peak_extra_columns(["BLS"]))
    n = int(rng.integers(2_000, 20_000))
    periods = np.sort(1.0 / np.linspace(0.025, 2.0, n))
    levels = int(rng.integers(3, 40))
    spectrum = {"power": np.round(rng.gamma(2.0, 1.0, n) * levels) / levels,
                "depth": rng.uniform(0, 1e-3, n), "depth_snr": rng.uniform(0, 9, n),
                "duration": rng.uniform(0.01, 0.2, n), "transit_time": rng.uniform(0, 5, n)}
    for want in (64, 512, 4096):
        cand = _as_candidates(periods, spectrum, want)
        forced = _as_candidates(periods, spectrum, want, cap=1)
        forced.written = max(forced.written, 2)  # over the cap: the power row decides
        _same(runner._candidate_row(prep, cand, settings, cfg),
              runner._candidate_row(prep, forced, settings, cfg))
