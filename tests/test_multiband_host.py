"""Multi-band BLS on the CPU: the oracle pinned (brute force, one-band limit,
invariances, phase scrambling), the input types and settings, and the host logic of
``bls_multiband_power`` with the device calls replaced by the oracle."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import multiband as omb
from oracle.bruteforce import joint_bruteforce
from synth import multiband_star

import cuperiod_amd as cup
from cuperiod_amd.core.errors import InsufficientDataError
from cuperiod_amd.methods import bls as bls_mod
from cuperiod_amd.methods import bls_mb
from cuperiod_amd.methods._bls_joint import joint_bands

BLS = cup.get_method("BLS")
#: The reference suite's small search (tests/test_multiband_bls_shared.py:26-29).
SETTINGS = cup.BLSSettings(
    min_period_days=1.0, max_period_days=6.0, n_durations=4, bins_per_duration=5,
    min_detections=10,
)


def _mblc(bands):
    return cup.MultiBandLightCurve.from_light_curves(
        {k: cup.LightCurve.from_arrays(t, y, e, domain="flux") for k, (t, y, e) in bands.items()}
    )


def _grid(mblc, settings=SETTINGS):
    time, value, error, _ = mblc.finite().stacked()
    return BLS.default_grid(cup.LightCurve.from_arrays(time, value, error, domain="flux"),
                            settings)


def _tiny_bands(seed):
    rng = np.random.default_rng(seed)
    times = [np.sort(rng.uniform(0, 20, n)) for n in (17, 23, 11)]
    vals = [1 + rng.normal(0, 0.01, t.size) for t in times]
    for v, t in zip(vals, times):
        v[np.mod(t, 3.1) < 0.2] -= 0.03
    errs = [0.01 * (1 + rng.uniform(0, 1, t.size)) for t in times]
    return times, vals, errs


# --- the oracle ------------------------------------------------------------------


@pytest.mark.parametrize("obj", [0, 1])
@pytest.mark.parametrize("seed", [3, 4])
def test_joint_oracle_equals_bruteforce(obj, seed):
    bands = omb.joint_bands(*_tiny_bands(seed))
    per, bd, db, width = omb.segment_spec(np.linspace(2.0, 4.0, 23), [0.1, 0.2, 0.35], 3)
    a = omb.joint_kernel(bands, per, bd, db, 3, width, obj)
    b = joint_bruteforce(bands, per, bin_duration=bd, dur_bins=db, oversample=3,
                         width=width, obj_flag=obj)
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_allclose(a[3], b[3], rtol=0, atol=1e-12 * np.abs(b[3]).max())
    np.testing.assert_allclose(a[4], b[4], rtol=1e-12)


def test_joint_oracle_overlapping_wrap_copy():
    # periods of a few bins: the wrap copy reads bins an earlier copy wrote
    bands = omb.joint_bands(*_tiny_bands(5))
    per, bd, db, width = omb.segment_spec(np.linspace(0.3, 0.5, 9), [0.1, 0.15], 4)
    a = omb.joint_kernel(bands, per, bd, db, 4, width, 0)
    b = joint_bruteforce(bands, per, bin_duration=bd, dur_bins=db, oversample=4,
                         width=width, obj_flag=0)
    np.testing.assert_array_equal(a[2], b[2])
    # a box over the wrapped copies holds the same points as one at the start: an exact
    # tie, which the oracle's rounded prefix sums may break the other way; the values
    # (snr objective = the reported power) agree
    fa = omb.finish(bands, list(a[3]), list(a[4]), a[0], a[1], a[2], per, bd, bands.floor)
    fb = omb.finish(bands, list(b[3]), list(b[4]), b[0], b[1], b[2], per, bd, bands.floor)
    np.testing.assert_allclose(fa["power"], fb["power"], rtol=1e-12)
    assert np.mean(a[0] == b[0]) >= 0.5


def test_joint_oracle_one_band_is_the_single_band_search():
    times, vals, errs = _tiny_bands(6)
    periods = np.linspace(2.0, 4.0, 41)
    durs = np.array([0.1, 0.2, 0.35])
    res = omb.joint_search(omb.joint_bands(times[:1], vals[:1], errs[:1]),
                           [(periods, durs)], 3)[0]
    single = oracle.bls_power(times[0], vals[0], errs[0], periods, durs, 3)
    np.testing.assert_array_equal(res["power"], np.clip(single["depth_snr"], 0, None) ** 2)
    # the same box, but where snr^2 rounds two boxes of distinct snr to one value (the
    # first of them is kept) -- boxes over the wrapped copies repeat boxes at the start
    same = res["transit_time"] == single["transit_time"]
    assert same.mean() >= 0.95
    np.testing.assert_array_equal(res["depth"][same], single["depth"][same])
    np.testing.assert_array_equal(res["duration"][same], single["duration"][same])


def test_joint_oracle_invariances():
    times, vals, errs = _tiny_bands(7)
    # dyadic times: a common shift by 2^10 d is exact
    times = [np.round(t * 2**20) / 2**20 for t in times]
    periods = np.linspace(2.0, 4.0, 31)
    seg = [(periods, np.array([0.1, 0.2, 0.35]))]
    a = omb.joint_search(omb.joint_bands(times, vals, errs), seg, 3)[0]
    b = omb.joint_search(omb.joint_bands([t + 1024.0 for t in times], vals, errs), seg, 3)[0]
    np.testing.assert_array_equal(a["power"], b["power"])
    np.testing.assert_allclose(b["transit_time"] - a["transit_time"], 1024.0, atol=1e-9)
    # a flux offset of one band is removed by its median
    c = omb.joint_search(
        omb.joint_bands(times, [vals[0] + 0.25, *vals[1:]], errs), seg, 3)[0]
    np.testing.assert_allclose(c["power"], a["power"], rtol=1e-9)


def test_phase_scrambled_band_no_longer_adds():
    # tests/test_multiband_bls_shared.py:193-228 on the oracle
    bands, period = multiband_star(11, n_bands=2, n_band=150, depth=0.03, chromatic=False)
    t1, y1, e1 = bands["b1"]
    mblc = _mblc(bands)
    grid = _grid(mblc)
    segs = bls_mod._grid_segments(grid, SETTINGS)
    aligned = omb.joint_search(
        omb.joint_bands([b[0] for b in bands.values()], [b[1] for b in bands.values()],
                        [b[2] for b in bands.values()]), segs, 5)
    scrambled = omb.joint_search(
        omb.joint_bands([bands["b0"][0], t1 + 0.375 * period], [bands["b0"][1], y1],
                        [bands["b0"][2], e1]), segs, 5)
    p_al = np.concatenate([s["power"] for s in aligned])
    p_sc = np.concatenate([s["power"] for s in scrambled])
    per = np.concatenate([p for p, _ in segs])
    i = int(np.argmax(p_al))
    assert per[i] == pytest.approx(period, rel=0.01)
    assert p_sc[i] < 0.75 * p_al[i]


# --- inputs and settings ---------------------------------------------------------------


def test_multiband_light_curve_checks():
    lc = cup.LightCurve.from_arrays([0, 1, 2], [1, 2, 3], domain="flux")
    with pytest.raises(ValueError):
        cup.MultiBandLightCurve.from_light_curves({})
    with pytest.raises(TypeError):
        cup.MultiBandLightCurve.from_light_curves({"g": lc, "r": ([0, 1], [1, 2])})
    with pytest.raises(ValueError, match="domains"):
        cup.MultiBandLightCurve.from_light_curves(
            {"g": lc, "r": cup.LightCurve.from_arrays([0, 1], [1, 2], domain="magnitude")})
    m = cup.MultiBandLightCurve.from_light_curves(
        {"g": lc, "r": cup.LightCurve.from_arrays([0.5, 1.5], [7, 8], [0.1, 0.2],
                                                  domain="flux")})
    t, v, e, b = m.stacked()
    np.testing.assert_array_equal(t, [0, 0.5, 1, 1.5, 2])
    np.testing.assert_array_equal(v, [1, 7, 2, 8, 3])
    assert e is None and list(b) == ["g", "r", "g", "r", "g"]
    assert m.band_names == ("g", "r") and m.n_bands == 2


def test_to_input_builds_multiband():
    g = cup.LightCurve.from_arrays([0, 1, 2], [10, 11, 12], domain="magnitude")
    r = cup.LightCurve.from_arrays([0, 1, 2], [1.0, 1.1, 1.2], domain="flux")
    with pytest.raises(ValueError, match="domain"):
        cup.to_input({"g": g, "r": r})
    m = cup.to_input({"g": g, "r": r}, domain="flux")
    assert isinstance(m, cup.MultiBandLightCurve)
    np.testing.assert_allclose(m.bands["g"].value, 10 ** (-0.4 * np.array([10, 11, 12.0])))


def test_resolve_ephemeris_and_scatter():
    assert bls_mb.resolve_ephemeris(SETTINGS) == "independent"
    shared = SETTINGS.model_copy(update={"mb_ephemeris": "shared"})
    assert bls_mb.resolve_ephemeris(shared) == "shared"
    assert bls_mb.resolve_ephemeris(shared, " Independent ") == "independent"
    assert bls_mb.resolve_ephemeris(object()) == "independent"
    with pytest.raises(ValueError, match="ephemeris"):
        bls_mb.resolve_ephemeris(SETTINGS, "joint")
    flux = np.array([1.0, 2.0, 3.0, 4.0, 100.0])
    assert bls_mb._noise_scale(flux) == pytest.approx(1.4826 * 1.0)
    assert bls_mb._noise_scale(np.array([1.0, 1.0, 1.0, 2.0])) == pytest.approx(
        np.std([1.0, 1.0, 1.0, 2.0]))
    assert bls_mb._noise_scale(np.ones(4)) == 1.0
    a = cup.LightCurve.from_arrays([0, 1, 2, 3, 4], flux, [1, 1, 1, 1, 1], domain="flux")
    b = cup.LightCurve.from_arrays([0, 1, 2, 3, 4], flux * 2, None, domain="flux")
    mixed = bls_mb._BandSet.select(_mblc_of({"a": a, "b": b}), 1)
    np.testing.assert_allclose(mixed.errors[0], 1.4826)
    np.testing.assert_allclose(mixed.errors[1], 2 * 1.4826)
    both = bls_mb._BandSet.select(_mblc_of({"a": a, "c": a}), 1)
    np.testing.assert_array_equal(both.errors[0], np.ones(5))


def _mblc_of(bands):
    return cup.MultiBandLightCurve.from_light_curves(bands)


def test_insufficient_data_is_refused():
    tiny = _mblc({"g": (np.array([0.0, 1.0]), np.array([1.0, 1.0]), np.array([1.0, 1.0]))})
    grid = cup.GridSpec(kind="period", values=np.array([1.5, 2.0]))
    with pytest.raises(InsufficientDataError, match="enough finite points"):
        bls_mb.bls_multiband_power(grid, tiny, SETTINGS, "hip", BLS)
    few = _mblc({k: v for k, v in multiband_star(1, n_band=4, n_bands=2)[0].items()})
    with pytest.raises(InsufficientDataError, match="min_detections"):
        bls_mb.bls_multiband_power(grid, few, SETTINGS, "hip", BLS)
    with pytest.raises(ValueError, match="ephemeris"):
        bls_mb.bls_multiband_power(grid, few, SETTINGS, "hip", BLS, ephemeris="joint")


# --- host logic with the oracle in place of the device ---------------------------------


def _oracle_segments_power(backend, jd, flux, err, segments, settings):
    out = []
    for periods, durations in segments:
        r = oracle.bls_power(jd, flux, err, periods, durations, settings.bins_per_duration,
                             objective=settings.objective)
        out.append({k: r[k] for k in ("period", "power", "depth", "depth_snr", "duration",
                                      "transit_time")})
    return out


def _oracle_joint_search(bands, segments, oversample, *, objective, backend, batch,
                         precision):
    ob = omb.JointBands(bands.tau, bands.yw, bands.ivar, bands.starts, bands.sum_y,
                        bands.sum_ivar, bands.floor, bands.t_min)
    return omb.joint_search(ob, segments, oversample, objective=objective)


def test_independent_stack_host_logic(monkeypatch):
    monkeypatch.setattr(bls_mod, "_segments_power", _oracle_segments_power)
    bands, _ = multiband_star(3)
    mblc = _mblc(bands)
    grid = _grid(mblc)
    pg = bls_mb.bls_multiband_power(grid, mblc, SETTINGS, "hip", BLS)
    segs = bls_mod._grid_segments(grid, SETTINGS)
    period, power, extras = omb.independent_stack(
        [b[0] for b in bands.values()], [b[1] for b in bands.values()],
        [b[2] for b in bands.values()], segs, 5)
    order = np.argsort(1.0 / period, kind="stable")
    np.testing.assert_array_equal(pg.power, power[order])
    for name in ("depth", "depth_snr", "duration", "t0", "band_index"):
        np.testing.assert_array_equal(pg.extras[name], extras[name][order], err_msg=name)
    assert pg.meta["mb_ephemeris"] == "independent"
    assert list(pg.extras) == ["depth", "depth_snr", "duration", "t0", "sde", "band_index"]
    assert pg.meta["bands"] == tuple(bands)


def test_shared_host_logic(monkeypatch):
    monkeypatch.setattr("cuperiod_amd.methods._bls_joint.joint_search", _oracle_joint_search)
    bands, period = multiband_star(12, n_bands=3, n_band=200, depth=0.03)
    mblc = _mblc(bands)
    grid = _grid(mblc)
    pg = bls_mb.bls_multiband_power(grid, mblc, SETTINGS, "hip", BLS, ephemeris="shared")
    i = int(np.argmax(pg.power))
    assert pg.period[i] == pytest.approx(period, rel=0.01)
    assert 0.0 < pg.extras["duration"][i] < 0.2 * period
    assert pg.extras["depth"][i] > 0.0
    assert pg.extras["depth_snr"][i] == pytest.approx(np.sqrt(pg.power[i]))
    assert pg.meta["bands"] == ("b0", "b1", "b2") and pg.meta["mb_ephemeris"] == "shared"
    # the host preparation is the oracle's
    prep = joint_bands([b[0] for b in bands.values()], [b[1] for b in bands.values()],
                       [b[2] for b in bands.values()])
    ref = omb.joint_bands([b[0] for b in bands.values()], [b[1] for b in bands.values()],
                          [b[2] for b in bands.values()])
    for name in ("tau", "yw", "ivar", "starts", "sum_y", "sum_ivar", "floor"):
        np.testing.assert_array_equal(getattr(prep, name), getattr(ref, name))
    assert prep.t_min == ref.t_min


def test_periodogram_routes_multiband(monkeypatch):
    monkeypatch.setattr("cuperiod_amd.methods._bls_joint.joint_search", _oracle_joint_search)
    monkeypatch.setattr(bls_mod, "_segments_power", _oracle_segments_power)
    monkeypatch.setattr(cup.get_method("BLS"), "resolve_backend_for", lambda b, s: "hip")
    bands, _ = multiband_star(13, n_bands=2)
    mblc = _mblc(bands)
    shared = SETTINGS.model_copy(update={"mb_ephemeris": "shared"})
    assert cup.periodogram(mblc, "BLS", settings=shared).meta["mb_ephemeris"] == "shared"
    assert cup.periodogram(mblc, "BLS", settings=SETTINGS).meta["mb_ephemeris"] == "independent"


def test_batch_rows_for_multiband_inputs(monkeypatch):
    from cuperiod_amd.batch import io as bio

    monkeypatch.setattr("cuperiod_amd.methods._bls_joint.joint_search", _oracle_joint_search)
    monkeypatch.setattr(bls_mod, "_segments_power", _oracle_segments_power)
    monkeypatch.setattr(cup.get_method("BLS"), "resolve_backend_for", lambda b, s: "hip")
    stars = [(f"s{i}", _mblc(multiband_star(40 + i, n_bands=3)[0])) for i in range(3)]
    stars.insert(1, ("empty", _mblc({"g": (np.array([0.0, 1.0]), np.array([1.0, 1.0]),
                                          np.array([1.0, 1.0]))})))
    settings = SETTINGS.model_copy(update={"mb_ephemeris": "shared"})
    res = cup.batch_periodograms(stars, "BLS", settings=settings)
    assert res.n_failed == 1 and res.errors[0][0] == "empty"
    extra = bio.peak_extra_columns(["BLS"], multiband=True)
    assert extra[-1] == "band_index"
    rows = iter(res.rows)
    for key, mblc in stars:
        if key == "empty":
            continue
        pg = cup.periodogram(mblc, "BLS", settings=settings)
        want = bio.periodogram_to_row(key, pg, n_best=10, store_raw=False, extra_keys=extra,
                                      source="")
        row = next(rows)
        assert row.keys() == want.keys()
        for k, v in want.items():
            assert v == row[k] or (isinstance(v, float) and np.isnan(v) and np.isnan(row[k])), k
