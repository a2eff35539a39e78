"""The learned path's refinement and fallback against the plain full-grid search.

Refinement searches a subset of the plain grid with the plain search's descriptors cut to
the windows, so its outputs at those periods are the plain outputs, and a refined curve
whose windows hold the plain best period reports it; a curve that falls back is the plain
search. These properties hold for any network weights.
"""

from __future__ import annotations

import numpy as np
import pytest

from cuperiod_amd import _lib
from cuperiod_amd.batch import engine
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.learned import proposer as lp
from cuperiod_amd.learned.network import init_parameters
from cuperiod_amd.learned.proposer import LearnedProposer
from cuperiod_amd.workloads import gapped_curves

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    print("case learned: 8 gapped curves x 30,000 frequencies", flush=True)
    curves, grids, _ = gapped_curves(8, seed=991, n_freq=30_000)
    settings = BLSSettings(n_durations=6)
    cs = engine.pack_curves(curves)
    prop = LearnedProposer(init_parameters(0), threshold=0.0, window_mass=1.0)
    prep = prop.prepare(cs, grids, settings)
    plain = engine.finish(prep.full, *engine.run_packed(prep.full), spectra=True)
    return settings, cs, prop, prep, plain


def test_window_batch_outputs_equal_the_full_search(batch):
    # A period's exact-sum outputs do not depend on what else is searched with it.
    settings, cs, prop, prep, _ = batch
    ranges = [(0, 700), (12_000, 13_500), (29_000, 30_000)]
    n = len(prep.grids)
    curve = np.repeat(np.arange(n), len(ranges))
    k0 = np.tile([a for a, _ in ranges], n)
    k1 = np.tile([b for _, b in ranges], n)
    wb = engine.window_batch(prep.full, curve, k0, k1)
    ctx = _lib.context(0)
    with ctx.peak_rounding(False):
        full_out, _ = engine.run_packed(prep.full)
        win_out, _ = engine.run_packed(wb)
    full_h = full_out.cpu().numpy()
    win_h = win_out.cpu().numpy()
    tab = wb.segment_table[:wb.n_segments]
    cols = np.arange(wb.n_out)
    r = np.searchsorted(tab["out_offset"], cols, side="right") - 1
    src = wb.source_col[r] + (cols - tab["out_offset"][r])
    np.testing.assert_array_equal(win_h[:, cols], full_h[:, src])
    assert wb.n_out == n * sum(b - a for a, b in ranges)


def test_stacked_features_are_finite(batch):
    settings, cs, prop, prep, _ = batch
    x = lp.coarse_features(prep).cpu().numpy()
    assert x.shape == (len(prep.grids), lp.L, 2)
    assert np.isfinite(x).all()
    assert prep.n_chunks >= 7 * len(prep.grids)  # ~100-day runs over 1,000 d with gaps


def test_refinement_never_beats_the_plain_maximum(batch):
    settings, cs, prop, prep, plain = batch
    r = prop.search(prep, settings)
    assert not r.fallback.any()
    assert np.all(r.best_power <= plain.best_power)
    for c in range(len(prep.grids)):
        periods = prep.full.plans[c].periods()
        p = np.nan_to_num(plain.spectra[c]["power"], nan=0.0, posinf=0.0, neginf=0.0)
        k = int(np.flatnonzero(periods == r.best_period[c])[0])
        # the same period of the plain search (its value may carry the other rounding:
        # each search rounds the near-maximum periods of what it searched)
        assert p[k] == pytest.approx(r.best_power[c], rel=1e-12)
        if r.best_power[c] == plain.best_power[c]:
            assert r.best_period[c] == plain.best_period[c]


def test_covered_best_period_is_found(batch):
    # windows placed on the plain best period (and elsewhere): the refined best period
    # is the plain one, bit for bit
    settings, cs, prop, prep, plain = batch
    n = len(prep.grids)
    k_best = np.array([int(np.flatnonzero(prep.full.plans[c].periods() == plain.best_period[c])[0])
                       for c in range(n)])
    curve = np.concatenate([np.arange(n), np.arange(n)])
    k0 = np.concatenate([np.maximum(k_best - 500, 0), np.full(n, 100)])
    k1 = np.concatenate([k_best + 500, np.full(n, 900)])
    c2, a, b = lp._merge_ranges(curve, k0, np.minimum(k1, prep.full.curve_count[curve]))
    wb = engine.window_batch(prep.full, c2, a, b)
    out, summary = engine.run_packed(wb)
    s = summary.cpu().numpy()
    best_col = wb.curve_base + s[:, 2].astype(np.int64)
    tab = wb.segment_table[:wb.n_segments]
    r = np.searchsorted(tab["out_offset"], best_col, side="right") - 1
    full_col = wb.source_col[r] + (best_col - tab["out_offset"][r])
    np.testing.assert_array_equal(engine.periods_of_columns(prep.full, full_col),
                                  plain.best_period)
    np.testing.assert_array_equal(s[:, 3], plain.best_power)


def test_full_fallback_is_the_plain_search(batch):
    settings, cs, prop, prep, plain = batch
    strict = LearnedProposer(init_parameters(0), threshold=2.0)
    r = strict.search(prep, settings)
    assert r.fallback.all()
    np.testing.assert_array_equal(r.best_period, plain.best_period)
    np.testing.assert_array_equal(r.best_power, plain.best_power)
