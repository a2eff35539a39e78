"""TLS on the CPU: the oracle pinned (brute force, template closed form, the
reference's known answers: transit recovery, unbiased depth, unit invariance, constant
flux), the default grid, and the host logic of ``tls_power`` with the device call
replaced by the oracle's sweep."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import tls as otls
from synth import inject_transit, nightly_times, template_matched_transit

import cuperiod_amd as cup
from cuperiod_amd.core import grid as grid_module
from cuperiod_amd.core.errors import GridSizeWarning, GridTooLargeError
from cuperiod_amd.methods import tls as tls_mod


def _prepared(seed=1, n=60):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0, 30, n))
    y = 1 + rng.normal(0, 0.01, n)
    y[np.mod(t, 3.3) < 0.2] -= 0.03
    dy = 0.01 * (1 + rng.uniform(0, 1, n))
    tt, _, w, w_total, yw = otls._prepare(t, y, dy)
    return tt - tt.min(), yw, w, w_total


# --- the oracle --------------------------------------------------------------------


@pytest.mark.parametrize(("n_bins", "dmin", "dmax", "nd"), [
    (32, 0.05, 0.3, 4), (16, 0.001, 0.01, 2), (24, 0.5, 0.9, 3), (64, 0.01, 0.1, 5)])
def test_oracle_sweep_equals_bruteforce(n_bins, dmin, dmax, nd):
    tau, yw, w, w_total = _prepared()
    periods = np.linspace(2.5, 4.0, 13)
    db = otls.duration_bins(n_bins, dmin, dmax, nd)
    tm = {wd: otls.limb_darkened_template(wd, 0.4, 0.3) for wd in db}
    a = otls.kernel(tau, yw, w, periods, n_bins, db, tm, w_total)
    b = otls.tls_bruteforce(tau, yw, w, periods, n_bins, db, tm, w_total)
    np.testing.assert_array_equal(a["duration"], b["duration"])
    np.testing.assert_array_equal(a["t0"], b["t0"])
    np.testing.assert_allclose(a["sr"], b["sr"], rtol=1e-12)
    np.testing.assert_allclose(a["depth"], b["depth"], rtol=1e-12)


def test_template_closed_form():
    g = otls.limb_darkened_template(21, 0.4, 0.3)
    assert g.size == 21 and g.max() == pytest.approx(1.0)
    assert g[10] == pytest.approx(1.0, abs=1e-6)
    np.testing.assert_allclose(g, g[::-1], rtol=0, atol=1e-15)
    x = (np.arange(21) + 0.5) / 21 - 0.5
    mu = np.sqrt(1 - (2 * x) ** 2)
    law = 1 - 0.4 * (1 - mu) - 0.3 * (1 - mu) ** 2
    np.testing.assert_allclose(g, law / law.max(), rtol=1e-15)
    np.testing.assert_array_equal(otls.limb_darkened_template(1, 0.4, 0.3), [1.0])
    np.testing.assert_array_equal(tls_mod.limb_darkened_template(21, 0.4, 0.3), g)


def test_oracle_recovers_a_planted_transit():
    t, flux, err = inject_transit(period=3.0, depth=0.02)
    periods = np.linspace(2.5, 3.5, 801)
    res = otls.tls_power(t, flux, err, periods)
    assert periods[int(np.argmax(res["sde"]))] == pytest.approx(3.0, rel=5e-3)
    assert res["sde"].max() > 7.0


@pytest.mark.parametrize("width_bins", [14, 26])
def test_oracle_depth_is_not_biased_by_the_duty_cycle(width_bins):
    # tests/test_tls.py:214-236
    assert width_bins in otls.duration_bins(256, 0.01, 0.10, 5)
    t, flux = template_matched_transit(width_bins)
    res = otls.tls_power(t, flux, None, np.array([3.0]))
    assert res["depth"][0] == pytest.approx(0.02, rel=0.015)
    assert res["duration"][0] == pytest.approx(3.0 * width_bins / 256.0)


@pytest.mark.parametrize("unit", [1e-15, 1e12])
def test_oracle_is_invariant_to_the_flux_unit(unit):
    t, flux, err = inject_transit()
    periods = np.linspace(2.0, 4.0, 300)
    ref = otls.tls_power(t, flux, err, periods)
    got = otls.tls_power(t, flux * unit, err * unit, periods)
    np.testing.assert_allclose(got["sde"], ref["sde"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(got["sr"], ref["sr"], rtol=1e-9, atol=0.0)
    np.testing.assert_allclose(got["depth"], ref["depth"] * unit, rtol=1e-9, atol=0.0)
    ref_nw = otls.tls_power(t, flux, None, periods)
    got_nw = otls.tls_power(t, flux * unit, None, periods)
    np.testing.assert_allclose(got_nw["sr"], ref_nw["sr"] * unit**2, rtol=1e-9, atol=0.0)


def test_oracle_constant_flux_and_no_dip():
    periods = 1.0 / np.linspace(1.0 / 60.0, 1.0, 600)[::-1]
    t = nightly_times(0)
    res = otls.tls_power(t, np.full(t.size, 12.345), None, periods)
    for key in ("sde", "sr", "depth", "duration"):
        assert np.count_nonzero(res[key]) == 0, key
    np.testing.assert_array_equal(res["t0"], np.full(periods.size, t.min()))


# --- settings, templates and the default grid ---------------------------------------------


def test_duration_bins_never_empty():
    db = tls_mod._duration_bins(cup.TLSSettings(
        n_phase_bins=16, duration_min_frac=0.001, duration_max_frac=0.01, n_durations=5))
    assert db and all(1 <= w < 16 for w in db)
    assert tls_mod._duration_bins(cup.TLSSettings()) == otls.duration_bins(256, 0.01, 0.1, 5)


def test_settings_bounds():
    with pytest.raises(ValueError, match="limb"):
        cup.TLSSettings(limb_dark_u1=0.8, limb_dark_u2=0.3)
    with pytest.raises(ValueError):
        cup.TLSSettings(duration_min_frac=0.2, duration_max_frac=0.1)


def test_grid_step_never_exceeds_the_narrowest_duration_budget():
    lc = cup.LightCurve.from_arrays(np.linspace(0.0, 100.0, 500), np.ones(500), domain="flux")
    method = cup.get_method("TLS")
    for settings in (
        cup.TLSSettings(),
        cup.TLSSettings(grid_duration_frac=0.2),
        cup.TLSSettings(grid_duration_frac=0.004),
        cup.TLSSettings(duration_min_frac=0.04, duration_max_frac=0.2),
    ):
        freq = method.default_grid(lc, settings).frequency
        step = float(np.median(np.diff(freq)))
        budget = min(settings.grid_duration_frac, 2.0 * settings.duration_min_frac)
        assert step == pytest.approx(budget / (settings.oversample * 100.0), rel=1e-6)


def test_grid_warns_and_refuses(monkeypatch):
    rng = np.random.default_rng(0)
    t = np.concatenate([[-3000.0], np.sort(rng.uniform(0.0, 30.0, 300))])
    lc = cup.LightCurve.from_arrays(t, np.ones(t.size), domain="flux")
    monkeypatch.setattr(grid_module, "AUTO_GRID_WARN_SIZE", 1000)
    with pytest.warns(GridSizeWarning, match="outlying"):
        cup.get_method("TLS").default_grid(lc, cup.TLSSettings())
    t = np.concatenate([np.linspace(0.0, 30.0, 300), [1e9]])
    lc = cup.LightCurve.from_arrays(t, np.ones(t.size), domain="flux")
    with pytest.raises(GridTooLargeError, match="TLS"):
        cup.get_method("TLS").default_grid(lc, cup.TLSSettings())


# --- host logic with the oracle's sweep in place of the device ------------------------------


def _oracle_device(tau, yw, w, periods, *, n_bins, dur_bins, templates, w_total, den_rel):
    return otls.kernel(tau, yw, w, periods, n_bins, dur_bins, templates, w_total, den_rel)


@pytest.mark.parametrize("errors", [True, False])
def test_tls_power_host_logic(monkeypatch, errors):
    monkeypatch.setattr(tls_mod, "_tls_hip", _oracle_device)
    t, flux, err = inject_transit()
    periods = np.linspace(2.0, 4.0, 150)
    dy = err if errors else None
    got = tls_mod.tls_power(t, flux, dy, periods, settings=cup.TLSSettings())
    ref = otls.tls_power(t, flux, dy, periods)
    for key in ("sde", "sr", "depth", "duration", "t0"):
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)


def test_tls_power_constant_flux_never_reaches_the_device(monkeypatch):
    def boom(*a, **k):
        raise AssertionError("device called")

    monkeypatch.setattr(tls_mod, "_tls_hip", boom)
    t = nightly_times(1)
    res = tls_mod.tls_power(t, np.full(t.size, 3.0), None, np.array([2.0, 3.0]),
                            settings=cup.TLSSettings())
    np.testing.assert_array_equal(res["t0"], np.full(2, t.min()))
    with pytest.raises(ValueError, match="backend"):
        tls_mod.tls_power(t, np.arange(t.size, dtype=float), None, np.array([2.0]),
                          settings=cup.TLSSettings(), backend="numpy")


def test_periodogram_routes_tls(monkeypatch):
    monkeypatch.setattr(tls_mod, "_tls_hip", _oracle_device)
    monkeypatch.setattr(cup.get_method("TLS"), "resolve_backend_for", lambda b, s: "hip")
    t, flux, err = inject_transit(period=3.0, depth=0.02)
    pg = cup.periodogram((t, flux, err), "TLS", domain="flux",
                         settings=cup.TLSSettings(min_period_days=2.0))
    assert pg.objective_sense == "max" and pg.method == "TLS"
    peak = pg.best_periods(1, alias_diverse=True)[0]
    assert peak.period == pytest.approx(3.0, rel=5e-3)
    assert {"depth", "duration", "t0", "sr"} <= set(peak.extra)


def test_batch_rows_equal_single_curve_rows(monkeypatch):
    from cuperiod_amd.batch import io as bio

    def oracle_curves(curves, *, n_bins, dur_bins, templates, den_rel):
        return [otls.kernel(tau, yw, w, per, n_bins, dur_bins, templates, w_total, den_rel)
                for tau, yw, w, w_total, per in curves]

    monkeypatch.setattr(tls_mod, "_tls_hip_curves", oracle_curves)
    monkeypatch.setattr(cup.get_method("TLS"), "resolve_backend_for", lambda b, s: "hip")
    items = []
    for seed in range(4):
        t, flux, err = inject_transit(seed=seed, period=2.0 + 0.4 * seed, n=400, span=30.0)
        items.append((f"c{seed}", cup.LightCurve.from_arrays(t, flux, err, domain="flux")))
    items.append(("few", cup.LightCurve.from_arrays([0.0, 1.0], [1.0, 2.0], domain="flux")))
    settings = cup.TLSSettings(min_period_days=1.0)
    res = cup.batch_periodograms(items, "TLS", settings=settings, chunk_size=2)
    assert res.n_failed == 1 and res.errors[0][0] == "few"
    extra = bio.peak_extra_columns(["TLS"])
    for row, (key, lc) in zip(res.rows, items):
        pg = cup.periodogram(lc, "TLS", settings=settings)
        want = bio.periodogram_to_row(key, pg, n_best=10, store_raw=False,
                                      extra_keys=extra, source="")
        assert row.keys() == want.keys()
        for k, v in want.items():
            same = (v == row[k]) or (isinstance(v, float) and np.isnan(v) and np.isnan(row[k]))
            assert same, (k, row[k], v)


def test_reciprocal_fold_guard_selects_the_division_bin():
    """The TLS kernel folds through tau * (1/P) and falls back to tau / P wherever q or
    ph * n_bins lies too close to an integer for the product to decide the bin; the bin
    chosen this way must equal the bin of the correctly rounded division everywhere,
    including points planted at exact multiples of the period and bin edges."""
    rng = np.random.default_rng(5)
    n_bins = 256
    periods = np.concatenate([rng.uniform(0.3, 15.0, 400), [0.5, 1.0, 3.0, 7.1, 10.0]])
    x = rng.uniform(0.0, 1200.0, 4000)
    for P in periods:
        k = rng.integers(0, 80, 600).astype(float)
        edge = (k + rng.integers(0, n_bins, 600) / n_bins) * P
        xs = np.concatenate([x, k * P, edge, np.nextafter(edge, 0), np.nextafter(edge, 1e9),
                             -x[:200]])
        q = xs / P
        want = np.clip((q - np.floor(q)) * n_bins, 0, n_bins - 1).astype(np.int64)
        qq = xs * (1.0 / P)
        ph = qq - np.floor(qq)
        r = ph * n_bins
        tq = np.abs(qq) * 2.0**-49
        redo = np.abs(r - np.rint(r)) <= (tq + 2.0**-50) * n_bins
        r[redo] = (q[redo] - np.floor(q[redo])) * n_bins
        got = np.clip(r, 0, n_bins - 1).astype(np.int64)
        np.testing.assert_array_equal(got, want)
