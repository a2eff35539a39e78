"""``batch_periodograms`` on the device: one engine call per chunk gives the rows of
the single-curve path bit for bit, and the oracle's best periods."""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from cuperiod_amd import BLSSettings, LightCurve, batch_periodograms, periodogram
from cuperiod_amd.batch import io as bio
from cuperiod_amd.methods.bls import BLSMethod, _grid_segments

pytestmark = pytest.mark.gpu


def _curves():
    rng = np.random.default_rng(77)
    out = []
    for i in range(11):
        n = int(rng.integers(150, 900))
        base = float(rng.uniform(20.0, 120.0))
        t = np.sort(rng.uniform(0.0, base, n)) + (2.45e6 if i % 3 == 0 else 0.0)
        y = 1.0 + rng.normal(0.0, 1e-3, n)
        p = float(rng.uniform(0.7, 6.0))
        y[np.mod(t, p) < 0.03 * p] -= 3e-3
        out.append((f"lc{i}", LightCurve(t, y, np.full(n, 1e-3), domain="flux")))
    out.insert(4, ("short", LightCurve([0.0, 1.0], [1.0, 1.0], [1.0, 1.0], domain="flux")))
    return out


def _same(a, b, sde_rtol=0.0):
    """Equal rows; ``sde_k`` within ``sde_rtol`` (device vs numpy mean/std)."""
    assert a.keys() == b.keys()
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, float) and math.isnan(x):
            assert isinstance(y, float) and math.isnan(y), k
        elif k.startswith("sde_") and sde_rtol:
            assert abs(x - y) <= sde_rtol * max(1.0, abs(y)), (k, x, y)
        else:
            assert x == y, (k, x, y)


def test_rows_equal_single_curve_path():
    items = _curves()
    res = batch_periodograms(items, "BLS", chunk_size=4)
    assert res.n_failed == 1 and res.errors[0][0] == "short"
    assert res.n_done == len(items) - 1
    extra = bio.peak_extra_columns(["BLS"])
    rows = iter(res.rows)
    for key, lc in items:
        if key == "short":
            continue
        pg = periodogram(lc, "BLS")
        expected = bio.periodogram_to_row(key, pg, n_best=10, store_raw=False,
                                          extra_keys=extra, source="")
        _same(next(rows), expected, sde_rtol=1e-12)


def test_raw_rows_equal_single_curve_path():
    items = [it for it in _curves() if it[0] != "short"][:4]
    res = batch_periodograms(items, "BLS", store_raw=True)
    extra = bio.peak_extra_columns(["BLS"])
    for row, (key, lc) in zip(res.rows, items):
        pg = periodogram(lc, "BLS")
        _same(row, bio.periodogram_to_row(key, pg, n_best=10, store_raw=True,
                                          extra_keys=extra, source=""))


def test_device_candidates_follow_the_host_rule():
    from cuperiod_amd.batch import engine
    from cuperiod_amd.core.config import BLSSettings as S
    from cuperiod_amd.workloads import survey_curves

    curves, grids, _ = survey_curves(6, seed=515, n_freq=40_000)
    settings = S(n_durations=6)
    pb = engine.pack(curves, grids, settings)
    out, summary = engine.run_packed(pb)
    res = engine.finish(pb, out, summary, spectra=True)
    for want in (5, 64, 512):
        pk = engine.peak_candidates(pb, out, want=want, cap=2 * want)
        for c in range(len(pb.plans)):
            s = np.nan_to_num(res.spectra[c]["power"][::-1], nan=0.0, posinf=0.0, neginf=0.0)
            inner = np.flatnonzero((s[1:-1] > s[:-2]) & (s[1:-1] >= s[2:])) + 1
            cands = list(inner)
            if s[0] >= s[1]:
                cands.append(0)
            if s[-1] > s[-2]:
                cands.append(s.size - 1)
            cands = np.sort(np.array(cands))
            assert pk.total[c] == cands.size
            got = pk.index[c]
            assert got.size == min(pk.written[c], 2 * want)
            assert pk.written[c] >= min(want, cands.size)
            thr = s[got].min()
            if pk.written[c] <= 2 * want:
                np.testing.assert_array_equal(got, cands[s[cands] >= thr])
            else:
                assert np.all(np.isin(got, cands))
            k = s.size - 1 - got
            np.testing.assert_array_equal(pk.values[c][:, 0], s[got])
            np.testing.assert_array_equal(pk.values[c][:, 1], res.spectra[c]["depth"][k])
            np.testing.assert_array_equal(pk.values[c][:, 4],
                                          res.spectra[c]["transit_time"][k])


def test_best_periods_match_oracle():
    items = [it for it in _curves() if it[0] != "short"][:5]
    settings = BLSSettings()
    res = batch_periodograms(items, "BLS")
    for row, (key, lc) in zip(res.rows, items):
        finite = lc.finite()
        grid = BLSMethod().default_grid(finite, settings)
        power = np.concatenate([
            oracle.bls_power(finite.time, finite.value, finite.error, p, d,
                             settings.bins_per_duration)["power"]
            for p, d in _grid_segments(grid, settings)
        ])
        periods = grid.values
        k = int(np.argmin(np.abs(periods - row["best_period"])))
        assert periods[k] == row["best_period"]
        power = np.nan_to_num(power, nan=0.0, posinf=0.0, neginf=0.0)
        # the oracle's own maximum, bit for bit (no tie tolerance)
        assert power[k] == power.max(), key
        assert np.count_nonzero(power == power.max()) == 1, key


def test_parquet_sink_resume(tmp_path):
    items = [it for it in _curves() if it[0] != "short"][:6]
    sink = tmp_path / "rows.parquet"
    first = batch_periodograms(items[:3], "BLS", sink=sink)
    second = batch_periodograms(items, "BLS", sink=sink, chunk_size=2)
    assert first.n_done == 3 and second.n_skipped == 3 and second.n_done == 3
    stored = bio.read_rows(sink)
    memory = batch_periodograms(items, "BLS").rows
    assert [r["key"] for r in stored] == [r["key"] for r in memory]
    for a, b in zip(stored, memory):
        _same(a, b)
