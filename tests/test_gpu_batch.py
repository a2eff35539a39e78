"""Batched engine vs single-curve engine and oracle (GPU).

The engine's bins are exact (64-bit fixed point), so a curve searched inside a batch
must give bit-identical outputs to the same curve searched alone on the same periods;
the device-regenerated parametric grids must be numpy's grids bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

import cuperiod_amd as cup
import oracle
from cuperiod_amd.batch.engine import LinspaceGrid, bls_batch
from cuperiod_amd.methods import _bls_core as core
from cuperiod_amd.methods.bls import BLSMethod, _grid_segments
from parity import KEYS, BoxInputs, compare

pytestmark = pytest.mark.gpu


def _curves(n_curves, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n_curves):
        n = int(np.round(np.exp(rng.uniform(np.log(100), np.log(2000)))))
        base = rng.uniform(50, 400)
        t = np.sort(rng.uniform(0, base, n)) + 2458000.0
        sig = np.exp(rng.uniform(np.log(5e-4), np.log(5e-3)))
        y = 1 + sig * rng.standard_normal(n)
        if c % 2 == 0:
            P = np.exp(rng.uniform(np.log(0.5), np.log(20)))
            q = rng.uniform(0.01, 0.05)
            y[np.mod(t - rng.uniform(0, P), P) < q * P] -= np.exp(rng.uniform(np.log(5e-4), np.log(1e-2)))
        out.append((t, y, np.full(n, sig)))
    return out


def test_linspace_batch_equals_single_curve_search():
    print("case batch linspace", flush=True)
    settings = cup.BLSSettings(n_durations=6)
    lcs = _curves(6, seed=1)
    grids = [LinspaceGrid(1.0 / (t.max() - t.min()), 2.0, 3000) for t, _, _ in lcs]
    res = bls_batch(lcs, grids, settings, spectra=True)
    for c, ((t, y, dy), g) in enumerate(zip(lcs, grids)):
        gs = g.as_gridspec()
        np.testing.assert_array_equal(res.plans[c].periods(), gs.period)
        segs = _grid_segments(gs, settings)
        assert [s.n_periods for s in res.plans[c].segments] == [p.size for p, _ in segs]
        single = core.bls_power_segments(t, y, dy, segs, settings.bins_per_duration)
        for i, k in enumerate(KEYS):
            ref = np.concatenate([getattr(s, k) for s in single])
            np.testing.assert_array_equal(res.spectra[c][k], ref, err_msg=f"{c}/{k}")
        # Against the oracle on the same periods.
        ref = {k: np.concatenate([oracle.bls_power(t, y, dy, p, d, 10)[k] for p, d in segs]) for k in KEYS}
        compare(res.spectra[c], ref, name=f"curve {c}",
                inputs=BoxInputs.for_specs(t, segs, settings.bins_per_duration))
        # Summary = the host assembly's SDE statistics and best period.
        pg = BLSMethod().power(gs, cup.LightCurve.from_arrays(t, y, dy, domain="flux"), settings, "hip")
        p = pg.power
        assert res.mean[c] == pytest.approx(float(p.mean()), rel=1e-12, abs=1e-300)
        assert res.std[c] == pytest.approx(float(p.std()), rel=1e-10)
        assert res.best_period[c] == pg.best_period()


def test_default_grid_batch_equals_single_curve_search():
    print("case batch default grid", flush=True)
    settings = cup.BLSSettings()
    lcs = _curves(4, seed=2)
    res = bls_batch(lcs, None, settings, spectra=True)
    for c, (t, y, dy) in enumerate(lcs):
        lc = cup.LightCurve.from_arrays(t, y, dy, domain="flux")
        grid = BLSMethod().default_grid(lc, settings)
        np.testing.assert_array_equal(res.plans[c].periods(), grid.period)
        pg = BLSMethod().power(grid, lc, settings, "hip")
        assert res.best_period[c] == pg.best_period()
        power = np.nan_to_num(res.spectra[c]["power"], neginf=0.0)
        np.testing.assert_array_equal(power[::-1], pg.power)


def test_survey_best_periods_identical_to_oracle():
    # 64 curves of the 200k-frequency survey batch against the oracle on every period:
    # power within 1e-6 (observed ~1e-13); the best period identical, with all seven
    # outputs bit-identical there; every period whose box differs from the oracle's is
    # proven to hold the same points (an exact tie the oracle's FP64 prefix sums order by
    # rounding). Flat-topped peaks (adjacent frequencies that fold the best box's points
    # identically) are exact ties too; the engine searches each segment's near-maximum
    # periods again with the reference's FP64 sequence, so it reports the oracle's sample.
    # Separate periods whose boxes hold the same points have equal exact statistics;
    # the oracle orders them by FP64 summation-order rounding (a few ulps) and numpy's
    # unstable sort of equal scores, and a plateau of such periods can be longer than the
    # engine's re-searched set: where the best periods differ, the two periods must be
    # proven to hold the same points, an exact tie the reference breaks by rounding.
    from cuperiod_amd.core.result import Periodogram
    from cuperiod_amd.workloads import survey_batch
    from parity import prove_cross_period_tie

    n_curves = 64
    print(f"case survey best periods, {n_curves} curves x 200,000 frequencies", flush=True)
    wl = survey_batch(n_curves, 200_000)
    settings = cup.BLSSettings(n_durations=wl.n_durations)
    identical = proven = sorted_ties = 0
    worst = 0.0
    for c0 in range(0, n_curves, 16):
        res = bls_batch(wl.curves[c0:c0 + 16], wl.grids[c0:c0 + 16], settings, spectra=True)
        for j, ((t, y, dy), g) in enumerate(zip(wl.curves[c0:c0 + 16], wl.grids[c0:c0 + 16])):
            c = c0 + j
            segs = _grid_segments(g.as_gridspec(), settings)
            parts = [oracle.bls_power(t, y, dy, p, d, 10) for p, d in segs]
            ref = {k: np.concatenate([p[k] for p in parts]) for k in KEYS}
            inputs = BoxInputs.for_specs(t, segs, settings.bins_per_duration)
            errs = compare(res.spectra[j], ref, name=f"survey curve {c}", inputs=inputs)
            period = np.concatenate([p["period"] for p in parts])
            rpow = np.nan_to_num(ref["power"], neginf=0.0)
            pg = Periodogram.from_spectrum(method="BLS", backend="oracle",
                                           frequency=1 / period, power=rpow,
                                           objective_sense="max", n_samples=1, baseline=1.0)
            gi = int(res.best_index[j])
            assert res.best_period[j] == period[gi]
            if res.best_period[j] == pg.best_period():
                identical += 1
            else:
                ri = int(np.flatnonzero(period == pg.best_period())[0])
                prove_cross_period_tie(ref, inputs, gi, ri, name=f"survey curve {c}")
                sorted_ties += 1
            proven += errs["ties_proven"]
            worst = max(worst, errs["power"])
        print(f"  curves {c0}..{c0 + 15}: best period identical on {identical}, proven "
              f"equal-point ties of the oracle's maximum on {sorted_ties}; {proven} "
              f"differing boxes proven ties; max rel power error {worst:.3g}", flush=True)
    assert identical + sorted_ties == n_curves


def test_non_finite_curve_leaves_its_neighbours_alone():
    # A curve with a non-finite flux value gets no box at any period (power -inf, the
    # other outputs 0); the curves packed beside it in the same launches, whose work
    # items finalize their periods together at the item's end, are unchanged.
    print("case batch non-finite", flush=True)
    settings = cup.BLSSettings(n_durations=6)
    lcs = _curves(4, seed=3)
    grids = [LinspaceGrid(1.0 / (t.max() - t.min()), 2.0, 3000) for t, _, _ in lcs]
    clean = bls_batch(lcs, grids, settings, spectra=True)
    t, y, dy = lcs[1]
    y = y.copy()
    y[len(y) // 2] = np.nan
    res = bls_batch([lcs[0], (t, y, dy), *lcs[2:]], grids, settings, spectra=True)
    for c in (0, 2, 3):
        for k in KEYS:
            np.testing.assert_array_equal(res.spectra[c][k], clean.spectra[c][k],
                                          err_msg=f"{c}/{k}")
    assert res.spectra[1]["power"].size == grids[1].as_gridspec().period.size
    assert np.all(res.spectra[1]["power"] == -np.inf)
    for k in KEYS[1:]:
        assert np.all(res.spectra[1][k] == 0.0), k
