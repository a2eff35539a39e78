"""Multi-band BLS on the device against the oracle: the joint-ephemeris search (both
objectives, both bin layouts: LDS and global scratch), the one-band limit, the
independent stack, what the statistic means (tests/test_multiband_bls_shared.py), and
the C ABI's multi-star batch."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from oracle import multiband as omb
from synth import multiband_star

import cuperiod_amd as cup
from cuperiod_amd import _lib
from cuperiod_amd.methods import bls as bls_mod
from cuperiod_amd.methods import bls_mb
from cuperiod_amd.methods._bls_joint import joint_bands

pytestmark = pytest.mark.gpu

BLS = cup.get_method("BLS")
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


def _oracle_shared(bands, grid, settings):
    """The oracle's joint search over the method's segments, in grid (period) order."""
    segs = bls_mod._grid_segments(grid, settings)
    ob = omb.joint_bands([b[0] for b in bands.values()], [b[1] for b in bands.values()],
                         [b[2] for b in bands.values()])
    found = omb.joint_search(ob, segs, settings.bins_per_duration,
                             objective=settings.objective)
    period = np.concatenate([p for p, _ in segs])
    out = {k: np.concatenate([f[k] for f in found]) for k in omb.JOINT_FIELDS}
    order = np.argsort(1.0 / period, kind="stable")
    return period[order], {k: v[order] for k, v in out.items()}


def _check_shared(pg, ref, *, same_min=0.999):
    period, r = ref
    np.testing.assert_array_equal(pg.period, period)
    scale = max(1.0, float(np.abs(r["power"]).max()))
    np.testing.assert_allclose(pg.power, r["power"], rtol=1e-9, atol=1e-9 * scale)
    same = np.isclose(pg.extras["t0"], r["transit_time"], rtol=0.0, atol=1e-9)
    same &= pg.extras["duration"] == r["duration"]
    assert same.mean() >= same_min, same.mean()
    # where the box is the same, every field agrees to round-off
    np.testing.assert_allclose(pg.power[same], r["power"][same], rtol=1e-12,
                               atol=1e-13 * scale)
    np.testing.assert_allclose(pg.extras["depth_snr"][same], r["depth_snr"][same],
                               rtol=1e-12, atol=1e-13 * np.sqrt(scale))
    np.testing.assert_array_equal(pg.extras["band_index"][same], r["band_index"][same])
    dscale = float(np.abs(r["depth"]).max()) or 1.0
    np.testing.assert_allclose(pg.extras["depth"][same], r["depth"][same], rtol=1e-9,
                               atol=1e-12 * dscale)
    assert pg.best_period() == period[int(np.argmax(r["power"]))]


@pytest.fixture(params=[0, 1], ids=["band-groups", "all-bands"])
def joint_kernel(request):
    """Both joint kernels: band groups (the default) and every band at once."""
    ctx = _lib.context(0)
    ctx.check(ctx.lib.cpb_set_joint_kernel(ctx.handle, request.param), "joint kernel")
    yield request.param
    ctx.check(ctx.lib.cpb_set_joint_kernel(ctx.handle, 0), "joint kernel")


@pytest.mark.parametrize("objective", ["snr", "likelihood"])
def test_shared_matches_oracle(objective, joint_kernel):
    bands, _ = multiband_star(7)
    mblc = _mblc(bands)
    grid = _grid(mblc)
    settings = SETTINGS.model_copy(update={"objective": objective})
    pg = bls_mb.bls_multiband_power(grid, mblc, settings, "hip", BLS, ephemeris="shared")
    assert pg.meta["mb_ephemeris"] == "shared"
    _check_shared(pg, _oracle_shared(bands, grid, settings))


@pytest.mark.parametrize("bins_per_duration", [2, 5])
def test_shared_many_durations_matches_oracle(bins_per_duration, joint_kernel):
    # 11-12 durations (the 512-thread band-group variants hold at most 8): over spans of
    # at most 403 bins (2 bins per duration) the 256-thread variant, over 1,006 bins (5)
    # the 1,024-thread one
    bands, _ = multiband_star(11)
    mblc = _mblc(bands)
    settings = SETTINGS.model_copy(update={"n_durations": 12,
                                           "bins_per_duration": bins_per_duration})
    grid = _grid(mblc, settings)
    pg = bls_mb.bls_multiband_power(grid, mblc, settings, "hip", BLS, ephemeris="shared")
    _check_shared(pg, _oracle_shared(bands, grid, settings))


@pytest.mark.parametrize("n_bands", [3, 6])
def test_shared_default_grid_matches_oracle(n_bands, joint_kernel):
    # default settings over a year: up to ~2000 phase bins per period, so 3 bands keep
    # their bins in LDS and 6 bands use the global-scratch layout
    bands, _ = multiband_star(21, n_bands=n_bands, n_band=120, span=365.0)
    mblc = _mblc(bands)
    settings = cup.BLSSettings(mb_ephemeris="shared")
    grid = _grid(mblc, settings)
    assert grid.values.size > 50_000
    pg = cup.periodogram(mblc, "BLS", settings=settings, grid=grid)
    _check_shared(pg, _oracle_shared(bands, grid, settings))


def test_one_band_shared_is_the_single_band_search(joint_kernel):
    bands, _ = multiband_star(9, n_bands=1, n_band=120)
    mblc = _mblc(bands)
    grid = _grid(mblc)
    shared = bls_mb.bls_multiband_power(grid, mblc, SETTINGS, "hip", BLS, ephemeris="shared")
    single = BLS.power(grid, mblc.bands["b0"], SETTINGS, "hip")
    expected = np.clip(single.extras["depth_snr"], 0.0, None) ** 2
    np.testing.assert_allclose(shared.power, expected, rtol=1e-12, atol=1e-12)
    # Boxes holding the same points tie exactly in the single-band engine (integer sums)
    # and in the all-bands joint kernel; the band-group kernel differences FP64 prefix
    # sums, as the reference does, so such ties break by rounding there. The powers
    # agree everywhere (above), so a different box is always a tie.
    same = np.isclose(shared.extras["t0"], single.extras["t0"], rtol=0.0, atol=1e-9)
    assert same.mean() >= (0.999 if joint_kernel == 1 else 0.99)


def test_independent_stack_matches_oracle():
    bands, _ = multiband_star(3)
    mblc = _mblc(bands)
    grid = _grid(mblc)
    pg = bls_mb.bls_multiband_power(grid, mblc, SETTINGS, "hip", BLS)
    segs = bls_mod._grid_segments(grid, SETTINGS)
    period, power, extras = omb.independent_stack(
        [b[0] for b in bands.values()], [b[1] for b in bands.values()],
        [b[2] for b in bands.values()], segs, SETTINGS.bins_per_duration)
    order = np.argsort(1.0 / period, kind="stable")
    np.testing.assert_allclose(pg.power, power[order], rtol=1e-10, atol=1e-12)
    same = np.isclose(pg.extras["t0"], extras["t0"][order], rtol=0.0, atol=1e-9)
    assert same.mean() >= 0.999
    np.testing.assert_array_equal(pg.extras["band_index"][same],
                                  extras["band_index"][order][same])
    assert pg.meta["mb_ephemeris"] == "independent"


def test_phase_scrambled_bands_no_longer_add_up():
    bands, period = multiband_star(11, n_bands=2, n_band=150, depth=0.03, chromatic=False)
    aligned = _mblc(bands)
    t1, y1, e1 = bands["b1"]
    scrambled = _mblc({"b0": bands["b0"], "b1": (t1 + 0.375 * period, y1, e1)})
    grid = _grid(aligned)
    i_al = bls_mb.bls_multiband_power(grid, aligned, SETTINGS, "hip", BLS)
    i_sc = bls_mb.bls_multiband_power(grid, scrambled, SETTINGS, "hip", BLS)
    s_al = bls_mb.bls_multiband_power(grid, aligned, SETTINGS, "hip", BLS, ephemeris="shared")
    s_sc = bls_mb.bls_multiband_power(grid, scrambled, SETTINGS, "hip", BLS,
                                      ephemeris="shared")
    assert i_sc.power.max() == pytest.approx(i_al.power.max(), rel=1e-6)
    assert np.median(np.abs(i_sc.power - i_al.power)) < 1e-9
    i = int(np.argmax(s_al.power))
    assert s_al.power[i] == pytest.approx(i_al.power[i], rel=0.05)
    assert s_al.best_period() == pytest.approx(period, rel=0.01)
    single = [BLS.power(grid, aligned.bands[b], SETTINGS, "hip").power.max() ** 2
              for b in ("b0", "b1")]
    assert s_sc.power[i] < 0.75 * s_al.power[i]
    assert s_sc.power[i] < 1.3 * max(single)


def test_weak_multiband_transit_is_recovered_by_the_shared_box():
    # tests/test_multiband_bls_shared.py:231-246: six sparse bands (45 points each), a
    # 1.5% transit in 0.5% noise
    n_stars = 30
    hits = {"independent": 0, "shared": 0}
    for seed in range(n_stars):
        bands, period = multiband_star(seed, depth=0.015)
        mblc = _mblc(bands)
        grid = _grid(mblc)
        for mode in hits:
            pg = bls_mb.bls_multiband_power(grid, mblc, SETTINGS, "hip", BLS, ephemeris=mode)
            hits[mode] += abs(pg.best_period() / period - 1.0) < 0.01
    assert hits["shared"] >= 0.9 * n_stars
    assert hits["shared"] - hits["independent"] >= 0.2 * n_stars


def test_shared_geometry_is_one_box_for_every_band():
    bands, period = multiband_star(12, n_bands=3, n_band=200, depth=0.03)
    mblc = _mblc(bands)
    pg = bls_mb.bls_multiband_power(_grid(mblc), mblc, SETTINGS, "hip", BLS,
                                    ephemeris="shared")
    i = int(np.argmax(pg.power))
    assert pg.period[i] == pytest.approx(period, rel=0.01)
    assert 0.0 < pg.extras["duration"][i] < 0.2 * period
    assert pg.extras["depth"][i] > 0.0
    assert pg.extras["depth_snr"][i] == pytest.approx(np.sqrt(pg.power[i]))
    assert pg.meta["bands"] == ("b0", "b1", "b2")


def test_settings_select_the_shared_ephemeris():
    bands, _ = multiband_star(13, n_bands=2)
    mblc = _mblc(bands)
    settings = SETTINGS.model_copy(update={"mb_ephemeris": "shared"})
    assert cup.periodogram(mblc, "BLS", settings=settings).meta["mb_ephemeris"] == "shared"
    default = cup.periodogram(mblc, "BLS", settings=SETTINGS)
    assert default.meta["mb_ephemeris"] == "independent"


def _abi_joint(stars_bands, periods, durations, oversample, obj_flag):
    """Several stars in one ``cpb_bls_joint_search`` call over one explicit segment
    each: the (6, n) output per star."""
    import torch

    from cuperiod_amd.methods._bls_core import _search_spec

    ctx = _lib.context(0)
    device = torch.device("cuda", ctx.device)
    prepared = [joint_bands(*zip(*[(t, y, e) for t, y, e in sb])) for sb in stars_bands]
    spec = _search_spec(periods, durations, oversample)
    taus, yws, ivs = [], [], []
    band_desc = (_lib.Band * sum(p.n_bands for p in prepared))()
    star_desc = (_lib.Star * len(prepared))()
    segs = (_lib.Segment * len(prepared))()
    point, b = 0, 0
    n = int(spec.periods.size)
    for s, p in enumerate(prepared):
        star_desc[s].band_offset = b
        star_desc[s].n_bands = p.n_bands
        star_desc[s].t_min = p.t_min
        for k in range(p.n_bands):
            band_desc[b].point_offset = point + int(p.starts[k])
            band_desc[b].n_points = int(p.starts[k + 1] - p.starts[k])
            band_desc[b].ivar_floor = float(p.floor[k])
            b += 1
        point += int(p.tau.size)
        taus.append(p.tau)
        yws.append(p.yw)
        ivs.append(p.ivar)
        sg = segs[s]
        sg.bin_duration = spec.bin_duration
        sg.period_offset = 0
        sg.n_periods = n
        sg.out_offset = s * n
        sg.width = spec.width
        sg.dur_offset = 0
        sg.n_dur = len(spec.dur_bins)
        sg.curve = s
        sg.mode = 0
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    tau, yw, iv = dev(np.concatenate(taus)), dev(np.concatenate(yws)), dev(np.concatenate(ivs))
    per = dev(spec.periods)
    dur = dev(np.asarray(spec.dur_bins, dtype=np.int32))
    out = torch.empty((_lib.N_JOINT_OUTPUTS, n * len(prepared)), dtype=torch.float64,
                      device=device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    with ctx.lock:
        ctx.check(ctx.lib.cpb_bls_joint_search(
            ctx.handle, stream, vp(tau), vp(yw), vp(iv), point, band_desc, b, star_desc,
            len(prepared), vp(per), vp(dur), segs, len(prepared), oversample, obj_flag,
            vp(out), n * len(prepared)), "joint search")
        host = out.cpu().numpy()
    return [host[:, s * n:(s + 1) * n] for s in range(len(prepared))]


def test_c_abi_stars_batch_equals_single_star_calls(joint_kernel):
    stars = []
    for seed, k in ((31, 2), (32, 5), (33, 1)):
        bands, _ = multiband_star(seed, n_bands=k, n_band=60)
        stars.append(list(bands.values()))
    periods = np.linspace(1.2, 3.0, 3001)
    durations = np.geomspace(0.02, 0.2, 5)
    for obj in (0, 1):
        together = _abi_joint(stars, periods, durations, 5, obj)
        for s, star in enumerate(stars):
            alone = _abi_joint([star], periods, durations, 5, obj)[0]
            np.testing.assert_array_equal(together[s], alone)
        # and the oracle's joint search on the same inputs
        for s, star in enumerate(stars):
            ob = omb.joint_bands(*zip(*star))
            ref = omb.joint_search(ob, [(periods, durations)], 5,
                                   objective="snr" if obj == 0 else "likelihood")[0]
            # sparse bands on a dense grid: many boxes hold the same points (exact ties
            # that the oracle's rounded prefix sums break either way); the power agrees
            # everywhere
            same = np.isclose(together[s][4], ref["transit_time"], rtol=0, atol=1e-9)
            assert same.mean() >= 0.99
            np.testing.assert_allclose(together[s][0], ref["power"], rtol=1e-9,
                                       atol=1e-9 * max(1.0, ref["power"].max()))


def test_oracle_single_band_sanity():
    # the single-band oracle the independent stack leans on, on one band of the star
    bands, _ = multiband_star(3, n_bands=1, n_band=80)
    t, y, e = bands["b0"]
    r = oracle.bls_power(t, y, e, np.linspace(1.5, 4.5, 200), [0.05, 0.1], 5)
    got = cup.bls_power(t, y, e, np.linspace(1.5, 4.5, 200), [0.05, 0.1], 5, backend="hip")
    np.testing.assert_allclose(got.power, r["power"], rtol=1e-10, atol=1e-12)


def test_batch_rows_equal_single_star_rows():
    from cuperiod_amd.batch import io as bio

    stars = [(f"s{i}", _mblc(multiband_star(50 + i, n_bands=3 + i % 2)[0])) for i in range(4)]
    single = [("lc", cup.LightCurve.from_arrays(*multiband_star(60, n_bands=1)[0]["b0"],
                                               domain="flux"))]
    settings = SETTINGS.model_copy(update={"mb_ephemeris": "shared"})
    res = cup.batch_periodograms(stars + single, "BLS", settings=settings, chunk_size=2)
    assert res.n_failed == 0 and len(res.rows) == len(stars) + 1
    extra = bio.peak_extra_columns(["BLS"], multiband=True)
    for row, (key, mblc) in zip(res.rows, stars):
        pg = cup.periodogram(mblc, "BLS", settings=settings)
        want = bio.periodogram_to_row(key, pg, n_best=10, store_raw=False, extra_keys=extra,
                                      source="")
        assert row.keys() == want.keys()
        for k, v in want.items():
            assert v == row[k] or (isinstance(v, float) and np.isnan(v) and np.isnan(row[k])), k
    assert res.rows[-1]["key"] == "lc" and np.isnan(res.rows[-1]["band_index_1"])
