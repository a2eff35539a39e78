"""HIP engine vs the FP64 oracle (parity bar in tests/parity.py).

Every case calls the engine through the C ABI (cpb_bls_search via ``bls_power`` /
``bls_power_segments``) and the oracle (the C++ restatement of the reference's numba
kernel) on the same seeded inputs.
"""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import oracle
from cuperiod_amd.methods import _bls_core as core
from cuperiod_amd.methods.bls import BLSMethod, _grid_segments
import cuperiod_amd as cup
from parity import KEYS, BoxInputs, compare
from synth import adversarial_cases, synthetic_eclipser, transit_curve, two_runs

pytestmark = pytest.mark.gpu


def _hip_power(t, y, dy, periods, durations, oversample, objective="snr"):
    out = core.bls_power(t, y, dy, periods, durations, oversample, objective=objective)
    return {k: getattr(out, k) for k in KEYS}


def _say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("objective", ["snr", "likelihood"])
def test_eclipser_matches_oracle(objective):
    _say(f"case eclipser objective={objective}")
    t, flux, err = synthetic_eclipser()
    periods = np.linspace(1.5, 4.0, 400)
    durations = np.array([0.05, 0.1, 0.2])
    got = _hip_power(t, flux, err, periods, durations, 10, objective)
    ref = oracle.bls_power(t, flux, err, periods, durations, 10, objective=objective)
    errs = compare(got, ref, name=f"eclipser/{objective}",
                   inputs=BoxInputs.for_specs(t, [(periods, durations)], 10))
    _say(f"  errors {errs}")
    assert errs["power"] < 1e-9


@pytest.mark.parametrize("obj_flag", [0, 1])
def test_adversarial_cases_match_oracle(obj_flag):
    # The reference's screen-exactness suite (tests/test_bls.py:374-424): offsets,
    # scales from 1e-200 to 1e160, constant and quantised flux, wild and huge errors.
    periods = np.linspace(0.6, 9.0, 160)
    durations = np.array([0.02, 0.05, 0.11, 0.3])
    objective = "snr" if obj_flag == 0 else "likelihood"
    for name, t, y, dy in adversarial_cases():
        _say(f"case adversarial {name!r} obj={obj_flag}")
        with np.errstate(all="ignore"):
            ref = oracle.bls_power(t, y, dy, periods, durations, 3, objective=objective)
            got = _hip_power(t, y, dy, periods, durations, 3, objective)
        # Exact-tie data (constant / quantised flux, five points in 800 bins) has
        # boxes holding identical point sets at most periods; the oracle orders them
        # by FP64 rounding, the engine's exact sums by first-wins. compare() proves
        # that every differing box holds the same points as the oracle's.
        errs = compare(got, ref, name=name,
                       inputs=BoxInputs.for_specs(t, [(periods, durations)], 3))
        _say(f"  errors {errs}")


def test_default_grid_periodogram_matches_oracle():
    _say("case default-grid periodogram")
    t, flux, err = transit_curve(n=1500, seed=8)
    lc = cup.LightCurve.from_arrays(t, flux, err, domain="flux")
    settings = cup.BLSSettings()
    method = BLSMethod()
    grid = method.default_grid(lc, settings)
    pg = method.power(grid, lc, settings, "hip")
    segs = _grid_segments(grid, settings)
    parts = [oracle.bls_power(t, flux, err, p, d, settings.bins_per_duration) for p, d in segs]
    ref_power = np.nan_to_num(np.concatenate([p["power"] for p in parts]), neginf=0.0)
    ref_freq = 1.0 / np.concatenate([p["period"] for p in parts])
    order = np.argsort(ref_freq, kind="stable")
    np.testing.assert_allclose(pg.power, ref_power[order], rtol=1e-9, atol=1e-12)
    # identical best period (the segments' near-maximum periods carry the reference's
    # FP64 rounding on the engine too)
    assert pg.best_period() == float(1.0 / ref_freq[order][np.argmax(ref_power[order])])
    _say(f"  best period {pg.best_period()}")


def test_two_runs_reject_swallowing_boxes():
    # tests/test_bls.py:182-210: with clustered sampling one box can hold every point;
    # such boxes must be rejected, and power is offset/scale invariant.
    settings = cup.BLSSettings(objective="likelihood", min_period_days=20.0)
    method = BLSMethod()
    for seed in (1, 5):
        _say(f"case two-runs seed={seed}")
        t, flux, err = two_runs(seed)
        lc = cup.LightCurve.from_arrays(t, flux, err, domain="flux")
        grid = method.default_grid(lc, settings)
        base = method.power(grid, lc, settings, "hip")
        assert base.best_periods(1)[0].extra["depth"] < 0.1
        assert float(base.power.max()) < 1e3
        for f2, e2 in ((flux + 1000.0, err), (flux * 1e12, err * 1e12)):
            other = method.power(
                grid, cup.LightCurve.from_arrays(t, f2, e2, domain="flux"), settings, "hip"
            )
            np.testing.assert_allclose(other.power, base.power, rtol=1e-6, atol=1e-9)


def test_recovers_period_and_depth():
    # tests/test_bls.py:73-79.
    _say("case recovery")
    t, flux, err = synthetic_eclipser(period=2.5, depth=0.05)
    pg = cup.periodogram((t, flux, err), "BLS", domain=cup.Domain.FLUX)
    peak = pg.best_periods(1, alias_diverse=True)[0]
    assert pg.backend == "hip"
    assert peak.period == pytest.approx(2.5, rel=3e-3)
    assert peak.extra["depth"] == pytest.approx(0.05, abs=0.01)
    assert peak.extra["sde"] > 7.0


def test_wide_histogram_path_matches_oracle():
    # Fixed short durations over long periods: > 10,000 phase bins per period, beyond
    # one workgroup's LDS, so the engine bins in global memory.
    _say("case wide histograms")
    rng = np.random.default_rng(3)
    t = np.sort(rng.uniform(0.0, 400.0, 2000))
    flux = 1.0 + 1e-3 * rng.standard_normal(t.size)
    flux[np.mod(t - 3.0, 61.0) < 0.3] -= 3e-3
    err = np.full(t.size, 1e-3)
    periods = np.linspace(55.0, 70.0, 120)
    durations = np.array([0.05, 0.1, 0.3])
    got = _hip_power(t, flux, err, periods, durations, 10)
    ref = oracle.bls_power(t, flux, err, periods, durations, 10)
    errs = compare(got, ref, name="wide",
                   inputs=BoxInputs.for_specs(t, [(periods, durations)], 10))
    _say(f"  errors {errs}")


def test_mid_span_histogram_path_matches_oracle():
    # ~7,000 phase bins per period: above the register-chunk prefix (2,048 bins), still
    # within one workgroup's LDS on a 160 KiB-LDS device (slot layout in place).
    _say("case mid-span histograms")
    rng = np.random.default_rng(4)
    t = np.sort(rng.uniform(0.0, 300.0, 3000))
    flux = 1.0 + 1e-3 * rng.standard_normal(t.size)
    flux[np.mod(t - 1.0, 29.0) < 0.2] -= 3e-3
    err = np.full(t.size, 1e-3)
    periods = np.linspace(25.0, 35.0, 150)
    durations = np.array([0.05, 0.1, 0.2])
    for objective in ("snr", "likelihood"):
        got = _hip_power(t, flux, err, periods, durations, 10, objective)
        ref = oracle.bls_power(t, flux, err, periods, durations, 10, objective=objective)
        errs = compare(got, ref, name=f"mid-span/{objective}",
                       inputs=BoxInputs.for_specs(t, [(periods, durations)], 10))
        _say(f"  errors {errs}")


@pytest.mark.parametrize("n_dur", [7, 20])
def test_many_durations_across_launch_classes_match_oracle(n_dur):
    # 7 and 20 box widths (the 8- and 32-duration builds) over periods whose spans run
    # from ~300 to ~1,500 bins: the segment is cut into the 128-thread, 8-wave,
    # 7-workgroup and 6-workgroup launch classes
    _say(f"case {n_dur} durations across launch classes")
    rng = np.random.default_rng(30 + n_dur)
    t = np.sort(rng.uniform(0.0, 120.0, 1500))
    flux = 1.0 + 1e-3 * rng.standard_normal(t.size)
    flux[np.mod(t - 1.3, 6.1) < 0.15] -= 2.5e-3
    err = np.full(t.size, 1e-3)
    periods = np.linspace(3.0, 15.0, 400)
    durations = np.geomspace(0.03, 0.5, n_dur)
    spec = core._search_spec(periods, durations, 3)
    assert spec.width > 1400
    got = _hip_power(t, flux, err, periods, durations, 3)
    ref = oracle.bls_power(t, flux, err, periods, durations, 3)
    errs = compare(got, ref, name=f"durations-{n_dur}",
                   inputs=BoxInputs.for_specs(t, [(periods, durations)], 3))
    _say(f"  errors {errs}")


def test_long_periods_on_dense_sampling_match_oracle():
    # 2,500 points over 20 d searched at 2-60 d with 0.1 d bins (spans of up to ~600
    # bins, the 128-thread class): consecutive points advance the phase by under a bin,
    # the case whose fold walks the curve in blocks by lane pairs (as in the learned
    # proposer's coarse search of time chunks)
    _say("case long periods, dense sampling")
    rng = np.random.default_rng(12)
    t = np.sort(rng.uniform(0.0, 20.0, 2500))
    flux = 1.0 + 1e-3 * rng.standard_normal(t.size)
    flux[np.mod(t - 2.0, 7.3) < 0.4] -= 2e-3
    err = np.full(t.size, 1e-3) * (1.0 + 0.5 * rng.random(t.size))
    periods = np.geomspace(2.0, 60.0, 300)
    durations = np.array([0.3, 0.6, 1.2])
    assert float(np.diff(t).mean()) * (60.0 / 0.1) / 60.0 < 1.0
    for objective in ("snr", "likelihood"):
        got = _hip_power(t, flux, err, periods, durations, 3, objective)
        ref = oracle.bls_power(t, flux, err, periods, durations, 3, objective=objective)
        errs = compare(got, ref, name=f"dense/{objective}",
                       inputs=BoxInputs.for_specs(t, [(periods, durations)], 3))
        _say(f"  errors {errs}")


@pytest.mark.parametrize("objective", ["snr", "likelihood"])
def test_skewed_flux_and_wide_boxes_match_oracle(objective):
    # The FP32 screen's error bounds at their edges: skewed flux (median far from the
    # weighted mean, so |ybar| is large in the screen's units), boxes holding most of
    # the weight (high words above 2^24, rounded when converted to float), and a few
    # points of tiny weight (near-empty out-of-box sides for the likelihood).
    _say(f"case skewed flux, wide boxes objective={objective}")
    rng = np.random.default_rng(11)
    t = np.sort(rng.uniform(0.0, 40.0, 2500))
    flux = 1e-3 * rng.standard_normal(t.size)
    flux[rng.random(t.size) < 0.4] += 1.0
    flux[np.mod(t - 0.7, 3.1) < 1.2] -= 0.3
    err = np.full(t.size, 1e-3)
    err[rng.random(t.size) < 0.01] = 1e2
    periods = np.linspace(0.8, 4.0, 300)
    durations = np.array([0.2, 0.6, 1.2, 2.0])
    got = _hip_power(t, flux, err, periods, durations, 5, objective)
    ref = oracle.bls_power(t, flux, err, periods, durations, 5, objective=objective)
    errs = compare(got, ref, name=f"skewed/{objective}",
                   inputs=BoxInputs.for_specs(t, [(periods, durations)], 5))
    _say(f"  errors {errs}")


def test_segments_equal_per_segment_calls():
    # tests/test_bls_segments.py:42-61: the segmented call equals one call per
    # segment; an empty segment comes back empty, in place.
    _say("case segments")
    t, flux, err = transit_curve()
    lc = cup.LightCurve.from_arrays(t, flux, err, domain="flux")
    settings = cup.BLSSettings()
    grid = BLSMethod().default_grid(lc, settings)
    segments = _grid_segments(grid, settings)
    segments = [segments[0], (np.zeros(0), segments[0][1]), *segments[1:]]
    got = core.bls_power_segments(t, flux, err, segments, settings.bins_per_duration)
    assert len(got) == len(segments)
    for (periods, durations), power in zip(segments, got):
        if periods.size == 0:
            assert power.power.size == 0
            continue
        ref = core.bls_power(t, flux, err, periods, durations, settings.bins_per_duration)
        for name in ("period", *KEYS):
            np.testing.assert_array_equal(getattr(power, name), getattr(ref, name))
