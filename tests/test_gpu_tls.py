"""TLS on the device against the oracle (tests/test_tls.py's cases): the default grid,
edge configurations of bins and durations, with and without errors, the reference's
known answers, and the C ABI's multi-curve batch."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import tls as otls
from synth import inject_transit, nightly_times, template_matched_transit

import cuperiod_amd as cup
from cuperiod_amd import _lib
from cuperiod_amd.methods.tls import tls_power

pytestmark = pytest.mark.gpu

TLS = cup.get_method("TLS")


def _oracle(t, y, dy, periods, settings):
    return otls.tls_power(
        t, y, dy, periods, n_bins=settings.n_phase_bins, d_min=settings.duration_min_frac,
        d_max=settings.duration_max_frac, n_dur=settings.n_durations,
        u1=settings.limb_dark_u1, u2=settings.limb_dark_u2)


def _check(got, ref, *, same_min=0.999):
    same = (got["duration"] == ref["duration"]) & np.isclose(got["t0"], ref["t0"], rtol=0,
                                                             atol=1e-9)
    assert same.mean() >= same_min, same.mean()
    np.testing.assert_allclose(got["sr"][same], ref["sr"][same], rtol=1e-9, atol=0)
    np.testing.assert_allclose(got["depth"][same], ref["depth"][same], rtol=1e-9, atol=0)
    # a different window only where two windows score the same to round-off
    np.testing.assert_allclose(got["sr"], ref["sr"], rtol=1e-9, atol=0)
    np.testing.assert_allclose(got["sde"], ref["sde"], rtol=1e-6, atol=1e-6)
    assert int(np.argmax(got["sde"])) == int(np.argmax(ref["sde"]))


def test_default_grid_matches_oracle():
    t, flux, err = inject_transit(period=3.0, depth=0.02)
    settings = cup.TLSSettings(min_period_days=2.0)
    grid = TLS.default_grid(cup.LightCurve.from_arrays(t, flux, err, domain="flux"), settings)
    periods = grid.period
    _check(tls_power(t, flux, err, periods, settings=settings),
           _oracle(t, flux, err, periods, settings))


@pytest.mark.parametrize(
    "overrides",
    [
        {},
        {"n_phase_bins": 16},
        {"duration_min_frac": 0.001, "duration_max_frac": 0.002, "n_durations": 1},
        {"duration_min_frac": 0.5, "duration_max_frac": 0.99, "n_durations": 4},
        {"n_phase_bins": 512, "n_durations": 8},
        {"n_phase_bins": 4096, "duration_max_frac": 0.2},
    ],
)
@pytest.mark.parametrize("errors", [True, False])
def test_edge_configurations_match_oracle(overrides, errors):
    # a transit centred on the first timestamp straddles fold phase 0 (circular wrap)
    t_first = float(inject_transit()[0].min())
    t, flux, err = inject_transit(t0=t_first)
    settings = cup.TLSSettings(**overrides)
    periods = np.linspace(2.0, 4.0, 150)
    dy = err if errors else None
    _check(tls_power(t, flux, dy, periods, settings=settings),
           _oracle(t, flux, dy, periods, settings))


def test_recovers_transit_through_periodogram():
    t, flux, err = inject_transit(period=3.0, depth=0.02)
    pg = cup.periodogram((t, flux, err), "TLS", domain="flux",
                         settings=cup.TLSSettings(min_period_days=2.0))
    peak = pg.best_periods(1, alias_diverse=True)[0]
    assert peak.period == pytest.approx(3.0, rel=5e-3)
    assert peak.extra["depth"] > 0.0 and peak.power > 7.0
    assert 0.0 < peak.extra["duration"] < 3.0
    mag = -2.5 * np.log10(inject_transit(depth=0.03)[1])
    pgm = cup.periodogram((t, mag, 0.003 * np.ones_like(mag)), "TLS", domain="magnitude",
                          settings=cup.TLSSettings(min_period_days=2.0))
    assert pgm.best_periods(1, alias_diverse=True)[0].period == pytest.approx(3.0, rel=1e-2)


def test_t0_is_an_absolute_mid_transit_time():
    true_t0, period = 2458001.0, 3.0
    t, flux, err = inject_transit(period=period, t0=true_t0)
    periods = np.linspace(2.9, 3.1, 201)
    res = tls_power(t, flux, err, periods, settings=cup.TLSSettings())
    best = int(np.argmax(res["sde"]))
    p_best, t0 = float(periods[best]), float(res["t0"][best])
    assert t.min() <= t0 < t.min() + p_best
    wrapped = ((t0 - true_t0) / p_best + 0.5) % 1.0 - 0.5
    assert abs(wrapped) < 2.0 / 256
    shifted = tls_power(t + 1000.37, flux, err, periods, settings=cup.TLSSettings())
    assert int(np.argmax(shifted["sde"])) == best
    assert shifted["t0"][best] - t0 == pytest.approx(1000.37, abs=1e-6)


@pytest.mark.parametrize("width_bins", [14, 26])
def test_depth_is_not_biased_by_the_duty_cycle(width_bins):
    t, flux = template_matched_transit(width_bins)
    res = tls_power(t, flux, None, np.array([3.0]), settings=cup.TLSSettings())
    assert res["depth"][0] == pytest.approx(0.02, rel=0.015)
    assert res["duration"][0] == pytest.approx(3.0 * width_bins / 256.0)


@pytest.mark.parametrize(("level", "errors"), [(12.345, False), (1234.5, True), (0.0, False)])
def test_constant_flux_gives_an_all_zero_spectrum(level, errors):
    periods = 1.0 / np.linspace(1.0 / 60.0, 1.0, 600)[::-1]
    t = nightly_times(1) + 0.123
    rng = np.random.default_rng(100)
    dy = (abs(level) or 1.0) * rng.uniform(5e-4, 3e-3, t.size) if errors else None
    res = tls_power(t, np.full(t.size, level), dy, periods,
                    settings=cup.TLSSettings(min_period_days=1.0))
    for key in ("sde", "sr", "depth", "duration"):
        assert np.count_nonzero(res[key]) == 0, key
    np.testing.assert_array_equal(res["t0"], np.full(periods.size, t.min()))


def test_tiny_real_transit_is_still_searched():
    t = nightly_times(7, n=600)
    flux = np.full(t.size, 12.345)
    flux[((t - t[0]) % 4.3) < 0.15] -= 12.345 * 1e-12
    periods = np.linspace(1.0, 20.0, 1500)
    res = tls_power(t, flux, None, periods, settings=cup.TLSSettings())
    assert np.count_nonzero(res["sde"]) > 0
    assert periods[int(np.argmax(res["sde"]))] == pytest.approx(4.3, rel=2e-3)


@pytest.mark.parametrize("unit", [1e-15, 1e12])
def test_float64_is_invariant_to_the_flux_unit(unit):
    t, flux, err = inject_transit()
    periods = np.linspace(2.0, 4.0, 300)
    ref = tls_power(t, flux, err, periods, settings=cup.TLSSettings())
    got = tls_power(t, flux * unit, err * unit, periods, settings=cup.TLSSettings())
    np.testing.assert_allclose(got["sde"], ref["sde"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(got["sr"], ref["sr"], rtol=1e-9, atol=0.0)
    np.testing.assert_allclose(got["depth"], ref["depth"] * unit, rtol=1e-9, atol=0.0)


def test_c_abi_curves_batch_equals_single_curve_calls():
    import torch

    from cuperiod_amd.methods._bls_core import _device_of, _stream_handle

    settings = cup.TLSSettings()
    db = otls.duration_bins(256, 0.01, 0.1, 5)
    flat = np.concatenate([otls.limb_darkened_template(w, 0.4, 0.3) for w in db])
    curves = []
    for seed, period in ((3, 2.2), (4, 3.1), (5, 1.7)):
        t, flux, err = inject_transit(seed=seed, period=period, n=900 + 100 * seed)
        _, _, w, w_total, yw = otls._prepare(t, flux, err)
        curves.append((t - t.min(), yw, w, w_total))
    periods = np.linspace(1.5, 3.5, 777)
    ctx = _lib.context(0)
    device = _device_of(ctx)

    def run(group):
        tau = np.concatenate([c[0] for c in group])
        yw = np.concatenate([c[1] for c in group])
        w = np.concatenate([c[2] for c in group])
        desc = (_lib.TlsCurve * len(group))()
        off = 0
        for i, c in enumerate(group):
            desc[i].point_offset = off
            desc[i].n_points = c[0].size
            desc[i].period_offset = 0
            desc[i].n_periods = periods.size
            desc[i].out_offset = i * periods.size
            desc[i].w_total = c[3]
            off += c[0].size
        d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
        tau_d, yw_d, w_d, per_d = d(tau), d(yw), d(w), d(periods)
        n_out = periods.size * len(group)
        out = torch.empty((_lib.N_TLS_OUTPUTS, n_out), dtype=torch.float64, device=device)
        widths = (ctypes.c_int32 * len(db))(*db)
        tm = np.ascontiguousarray(flat)
        vp = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
        with ctx.lock:
            ctx.check(ctx.lib.cpb_tls_search(
                ctx.handle, _stream_handle(device), vp(tau_d), vp(yw_d), vp(w_d), tau.size,
                desc, len(group), vp(per_d), 256, widths, len(db),
                tm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1e-9, vp(out), n_out),
                "tls")
            return out.cpu().numpy()

    together = run(curves)
    for i, c in enumerate(curves):
        alone = run([c])
        np.testing.assert_array_equal(together[:, i * periods.size:(i + 1) * periods.size],
                                      alone)
        ref = otls.kernel(c[0], c[1], c[2], periods, 256, db,
                          {w: otls.limb_darkened_template(w, 0.4, 0.3) for w in db}, c[3])
        np.testing.assert_allclose(alone[0], ref["sr"], rtol=1e-9, atol=0)
    assert settings.n_phase_bins == 256


def test_batch_equals_single_curve_calls():
    from cuperiod_amd.methods.tls import tls_power_batch

    settings = cup.TLSSettings()
    curves, periods = [], []
    for seed in range(5):
        t, flux, err = inject_transit(seed=seed, period=2.0 + 0.3 * seed, n=700 + 50 * seed)
        curves.append((t, flux, err if seed % 2 else None))
        periods.append(np.linspace(1.5, 4.0, 500 + 37 * seed))
    t = nightly_times(2)
    curves.append((t, np.full(t.size, 2.0), None))  # constant: never reaches the device
    periods.append(np.linspace(1.0, 3.0, 50))
    batch = tls_power_batch(curves, periods, settings=settings)
    for (t, y, dy), per, got in zip(curves, periods, batch):
        alone = tls_power(t, y, dy, per, settings=settings)
        for key in ("sde", "sr", "depth", "duration", "t0"):
            np.testing.assert_array_equal(got[key], alone[key], err_msg=key)


def test_batch_periodograms_rows_equal_single_curve_rows():
    from cuperiod_amd.batch import io as bio

    items = []
    for seed in range(5):
        t, flux, err = inject_transit(seed=seed, period=2.0 + 0.4 * seed, n=500, span=40.0)
        items.append((f"c{seed}", cup.LightCurve.from_arrays(t, flux, err, domain="flux")))
    settings = {"TLS": cup.TLSSettings(min_period_days=1.0)}
    res = cup.batch_periodograms(items, ["BLS", "TLS"], settings=settings, chunk_size=3)
    assert res.n_failed == 0
    extra = bio.peak_extra_columns(["BLS", "TLS"])
    tls_rows = [r for r in res.rows if r["method"] == "TLS"]
    assert len(tls_rows) == len(items)
    for row, (key, lc) in zip(tls_rows, items):
        pg = cup.periodogram(lc, "TLS", settings=settings["TLS"])
        want = bio.periodogram_to_row(key, pg, n_best=10, store_raw=False,
                                      extra_keys=extra, source="")
        assert row.keys() == want.keys()
        for k, v in want.items():
            same = (v == row[k]) or (isinstance(v, float) and np.isnan(v) and np.isnan(row[k]))
            assert same, (k, row[k], v)


def test_dense_sampling_matches_oracle():
    # 20,000 points over 27 d (the TLS survey's curves): at the longer periods a wave's 64
    # consecutive points share a few of the 256 phase bins, so the kernel's run sums (one
    # atomic pair per run of equal bins) carry most of the fold.
    from cuperiod_amd.workloads import tls_survey

    curves, _ = tls_survey(n_curves=1)
    t, flux, err = curves[0]
    settings = cup.TLSSettings()
    grid = TLS.default_grid(cup.LightCurve.from_arrays(t, flux, err, domain="flux"), settings)
    periods = np.ascontiguousarray(grid.period[::13])
    _check(tls_power(t, flux, err, periods, settings=settings),
           _oracle(t, flux, err, periods, settings))
